// The absorbing atom's Cartesian cross-section tensors on the device
// (Tenseur_car, tensor.f90:15-946, with tens_ab :947-1345): the electric
// dipole-dipole (secdd), dipole-quadrupole (secdq) and quadrupole-quadrupole
// (secqq) tensors of every energy point from the atom's amplitude products
// Taull and its radial integrals rof.  The branch of an FDM absorption
// calculation: no TDDFT, optic, NRIXS, Green's-function or FDM_comp output,
// no singular solution, no spin-orbit, one final-state spin, spherical
// potential, real harmonics.
//
// Everything that does not depend on the energy is planned on the host at
// creation: the core states of the edge (coef_init, general.f90:774), the
// Gaunt coefficients (here by quadrature of three real harmonics on a
// Gauss-Legendre x uniform-azimuth product grid, exact for the degrees
// involved), the product terms of tens_ab with their coefficients Cg, the
// Cartesian-spherical weights clm (tensor.f90:240-275), the spreading of
// each independent element to the elements equal by symmetry (the msym
// tables, tensor.f90:529-700) and the frame rotation (:781-806).  One
// workgroup per energy point then evaluates, in LDS, the spherical
// components, the independent Cartesian elements, the spread tensors and
// their rotation, every sum in the reference's order.
#include <algorithm>
#include <cmath>
#include <complex>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "solver.h"
#include "tensor.h"
#include "ylm.h"

#pragma clang fp contract(off)

namespace fdmgpu {
namespace {

typedef std::complex<double> cd;
constexpr double kPi = 3.1415926535897932384626433832;

// ---- host: Gaunt coefficients of real harmonics by quadrature ----

void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; i++) {
    double z = std::cos(kPi * (i + 0.75) / (n + 0.5)), dp = 1;
    for (int it = 0; it < 100; it++) {
      double p0 = 1, p1 = z;
      for (int k = 2; k <= n; k++) {
        const double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1);
      const double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    x[i] = z;
    w[i] = 2 / ((1 - z * z) * dp * dp);
  }
}

// G(l1 m1, l2 m2, l3 m3) = integral of Y_l1m1 Y_l2m2 Y_l3m3 over the sphere
struct GauntTable {
  static constexpr int LMAX = 3, N = (LMAX + 1) * (LMAX + 1);
  std::vector<double> g;
  GauntTable() : g((size_t)N * N * N, 0.0) {
    const int nt = 12, np = 24;  // exact for polynomial degree < 24 in cos(theta), azimuthal order < 24
    std::vector<double> x, w;
    gauss_legendre(nt, x, w);
    std::vector<double> y(N);
    for (int i = 0; i < nt; i++) {
      const double ct = x[i], st = std::sqrt(1 - ct * ct);
      for (int j = 0; j < np; j++) {
        const double ph = 2 * kPi * j / np, wt = w[i] * 2 * kPi / np;
        real_ylm(LMAX, st * std::cos(ph), st * std::sin(ph), ct, y.data());
        for (int a = 0; a < N; a++)
          for (int b = 0; b < N; b++) {
            const double ab = wt * y[a] * y[b];
            for (int c = 0; c < N; c++) g[((size_t)a * N + b) * N + c] += ab * y[c];
          }
      }
    }
  }
  double r(int l1, int m1, int l2, int m2, int l3, int m3) const {
    return g[((size_t)(l1 * l1 + l1 + m1) * N + (l2 * l2 + l2 + m2)) * N + (l3 * l3 + l3 + m3)];
  }
  // Y_lm and Y_lomo real, Y_limi the complex harmonic written with the real
  // ones as Gaunt_xrc does (tensor.f90:1896-1907)
  cd xrc(int l, int m, int lo, int mo, int li, int mi) const {
    const double s2 = std::sqrt(2.0);
    if (mi == 0) return cd(r(l, m, lo, mo, li, mi), 0.0);
    const double sg = (mi & 1) ? -1.0 : 1.0;
    if (mi > 0) return cd(sg * r(l, m, lo, mo, li, mi) / s2, sg * r(l, m, lo, mo, li, -mi) / s2);
    return cd(r(l, m, lo, mo, li, -mi) / s2, -r(l, m, lo, mo, li, mi) / s2);
  }
};

// ---- host: the core states of the edge(s), coef_init (general.f90:774-839) ----
struct CoreState {
  double c[2] = {0, 0};  // weight of the spin-up / spin-down component
  int m[2] = {1000, 1000};
};

std::vector<CoreState> core_states(int jseuil, int lseuil, int nbseuil, int &ninit1) {
  std::vector<CoreState> out;
  int is_init = (jseuil == 1 || jseuil == 3 || jseuil == 5 || jseuil == 7) ? 1 : -1;
  ninit1 = 0;
  for (int e = 1; e <= nbseuil; e++) {
    if (e == 2) is_init = -is_init;
    const double J = lseuil + 0.5 * is_init;
    const int nj = (int)std::lround(2 * J + 1);
    if (e == 1) ninit1 = nj;
    for (int i = 1; i <= nj; i++) {
      CoreState s;
      const double Jz = -J + i - 1;
      for (int isp = 0; isp < 2; isp++) {
        const double spin = isp == 0 ? 0.5 : -0.5;
        const int sg = isp == 0 ? is_init : 1;
        const int mi = (int)std::lround(Jz - spin);
        if (std::abs(mi) > lseuil) continue;
        const double c2 = (lseuil + 0.5 + 2 * spin * is_init * Jz) / (2 * lseuil + 1);
        if (c2 < 1e-6) continue;
        s.c[isp] = sg * std::sqrt(c2);
        s.m[isp] = mi;
      }
      out.push_back(s);
    }
  }
  return out;
}

// clm[irang][k][j][lm]: Cartesian -> spherical weights of the dipole (irang
// 1) and quadrupole (irang 2) operators, tensor.f90:240-275
void clm_table(double clm[3][3][3][16]) {
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++)
      for (int c = 0; c < 3; c++)
        for (int d = 0; d < 16; d++) clm[a][b][c][d] = 0;
  const double c1 = std::sqrt(4 * kPi / 3);
  for (int j = 0; j < 3; j++) {
    clm[1][0][j][3] = c1;  // x
    clm[1][1][j][1] = c1;  // y
    clm[1][2][j][2] = c1;  // z
  }
  const double c0 = std::sqrt(4 * kPi) / 3, c = std::sqrt(4 * kPi / 15), c3 = c / std::sqrt(3.0);
  clm[2][0][0][0] = c0;
  clm[2][0][0][6] = -c3;
  clm[2][0][0][8] = c;
  clm[2][0][1][4] = c;
  clm[2][0][2][7] = c;
  clm[2][1][1][0] = c0;
  clm[2][1][1][6] = -c3;
  clm[2][1][1][8] = -c;
  clm[2][1][2][5] = c;
  clm[2][2][2][0] = c0;
  clm[2][2][2][6] = 2 * c3;
  for (int i = 0; i < 3; i++)
    for (int j = i + 1; j < 3; j++)
      for (int lm = 0; lm < 16; lm++) clm[2][j][i][lm] = clm[2][i][j][lm];
}

__host__ __device__ inline int i3(int a, int b) { return a * 3 + b; }
__host__ __device__ inline int i4(int a, int b, int c) { return (a * 3 + b) * 3 + c; }
__host__ __device__ inline int i5(int a, int b, int c, int d) { return ((a * 3 + b) * 3 + c) * 3 + d; }

template <class T>
int to_device(const std::vector<T> &h, T **d, std::vector<void *> &owned) {
  FDM_HIP(hipMalloc((void **)d, sizeof(T) * std::max<size_t>(h.size(), 1)));
  owned.push_back(*d);
  if (!h.empty()) FDM_HIP(hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return 0;
}

// ---- device ----
struct z2 {
  double re, im;
};
__device__ inline z2 zmul(z2 a, z2 b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ inline z2 zconj(z2 a) { return {a.re, -a.im}; }
__device__ inline z2 zld(const double *p) { return {p[0], p[1]}; }

__global__ void __launch_bounds__(256) k_tensor(TensDev D, TensPoints P, int64_t toff, int ld) {
  extern __shared__ z2 sm[];
  const int g = blockIdx.x, ni = D.ninitr;
  const int na = max(D.ncomp, TENS_NOUT) * ni;
  z2 *A = sm, *B = sm + na;
  const double *rof = P.rof[g], *tau = P.taull[g] + 2 * toff;
  // spherical components Tens_lm (tens_ab): per (component, initr)
  for (int w = threadIdx.x; w < D.ncomp * ni; w += blockDim.x) {
    z2 acc = {0, 0};
    for (int q = D.term_ptr[w]; q < D.term_ptr[w + 1]; q++) {
      const TensTerm t = D.terms[q];
      const z2 r1 = zld(rof + 2 * t.r1), r2 = zld(rof + 2 * t.r2);
      const z2 t12 = zld(tau + 2 * ((int64_t)t.lm1 * ld + t.lm2)), t21 = zld(tau + 2 * ((int64_t)t.lm2 * ld + t.lm1));
      const z2 cfe = zmul(zmul(zconj(r1), r2), t12);
      const z2 cfs = zmul(zmul(zconj(r2), r1), t21);
      const z2 tr = {0.5 * (cfe.re - cfs.re), 0.5 * (cfe.im + cfs.im)};
      const z2 ct = zmul({t.cg_re, t.cg_im}, tr);
      acc = {acc.re + ct.re, acc.im + ct.im};
    }
    acc = {-acc.im, acc.re};  // times i (tensor.f90:1327)
    if (D.hydrogen) acc = {acc.re / 2, acc.im / 2};
    A[w] = acc;
  }
  __syncthreads();
  // independent Cartesian elements: sum of clm_s clm_i Tens_lm
  for (int w = threadIdx.x; w < D.nrep * ni; w += blockDim.x) {
    const int e = w / ni, k = w % ni;
    z2 acc = {0, 0};
    for (int q = D.rep_ptr[e]; q < D.rep_ptr[e + 1]; q++) {
      const double c = D.rep_w[q];
      const z2 v = A[D.rep_comp[q] * ni + k];
      acc = {acc.re + c * v.re, acc.im + c * v.im};
    }
    B[w] = acc;
  }
  __syncthreads();
  // spread to the elements equal by symmetry
  for (int w = threadIdx.x; w < TENS_NOUT * ni; w += blockDim.x) {
    const int o = w / ni, k = w % ni;
    const int *f = D.fill + 4 * o;
    z2 v = {0, 0};
    if (f[0] >= 0) {
      const z2 t = B[f[0] * ni + k];
      v = {f[1] * t.re, f[2] * t.im};
      if (f[3]) v.im = -v.im;
    }
    A[w] = v;
  }
  __syncthreads();
  // rotation R T R^T (rot_tensor_2/3/4, tensor.f90:2112-2210), output
  const double *R = D.rot;
  for (int w = threadIdx.x; w < TENS_NOUT * ni; w += blockDim.x) {
    const int o = w / ni, k = w % ni;
    z2 out = {0, 0};
    if (o < 9) {
      if (!D.out_dd) continue;
      const int i = o / 3, j = o % 3;
      for (int kk = 0; kk < 3; kk++) {
        z2 cmat = {0, 0};
        for (int l = 0; l < 3; l++) {
          const z2 v = A[i3(kk, l) * ni + k];
          cmat = {cmat.re + R[j * 3 + l] * v.re, cmat.im + R[j * 3 + l] * v.im};
        }
        out = {out.re + R[i * 3 + kk] * cmat.re, out.im + R[i * 3 + kk] * cmat.im};
      }
      double *d = P.dd[g] + 2 * (k * 9 + o);
      d[0] = out.re;
      d[1] = out.im;
    } else if (o < 36) {
      if (!D.out_dq) continue;
      const int p = o - 9, i = p / 9, j = (p / 3) % 3, kk = p % 3;
      for (int l = 0; l < 3; l++) {
        z2 cmas = {0, 0};
        for (int m = 0; m < 3; m++) {
          z2 cmat = {0, 0};
          for (int n = 0; n < 3; n++) {
            const z2 v = A[(9 + i4(l, m, n)) * ni + k];
            cmat = {cmat.re + R[kk * 3 + n] * v.re, cmat.im + R[kk * 3 + n] * v.im};
          }
          cmas = {cmas.re + R[j * 3 + m] * cmat.re, cmas.im + R[j * 3 + m] * cmat.im};
        }
        out = {out.re + R[i * 3 + l] * cmas.re, out.im + R[i * 3 + l] * cmas.im};
      }
      double *d = P.dq[g] + 2 * (k * 27 + p);
      d[0] = out.re;
      d[1] = out.im;
    } else {
      if (!D.out_qq) continue;
      const int p = o - 36, i = p / 27, j = (p / 9) % 3, kk = (p / 3) % 3, l = p % 3;
      for (int m = 0; m < 3; m++) {
        z2 cmar = {0, 0};
        for (int n = 0; n < 3; n++) {
          z2 cmas = {0, 0};
          for (int n1 = 0; n1 < 3; n1++) {
            z2 cmat = {0, 0};
            for (int q = 0; q < 3; q++) {
              const z2 v = A[(36 + i5(m, n, n1, q)) * ni + k];
              cmat = {cmat.re + R[l * 3 + q] * v.re, cmat.im + R[l * 3 + q] * v.im};
            }
            cmas = {cmas.re + R[kk * 3 + n1] * cmat.re, cmas.im + R[kk * 3 + n1] * cmat.im};
          }
          cmar = {cmar.re + R[j * 3 + n] * cmas.re, cmar.im + R[j * 3 + n] * cmas.im};
        }
        out = {out.re + R[i * 3 + m] * cmar.re, out.im + R[i * 3 + m] * cmar.im};
      }
      double *d = P.qq[g] + 2 * (k * 81 + p);
      d[0] = out.re;
      d[1] = out.im;
    }
  }
}

}  // namespace

TensorPlan::~TensorPlan() {
  for (void *p : owned) (void)hipFree(p);
}

int TensorPlan::init(const fdmgpu_tensor_tables *t, int maxp, hipStream_t s) {
  if (!t || maxp < 1 || maxp > TENS_MAXP || t->lseuil < 0 || t->lseuil > 3 || t->nbseuil < 1 || t->nbseuil > 2 ||
      t->lmax < 0 || t->lmax > 3 || !(t->e1e1 || t->e1e2 || t->e2e2) || !t->ldip || !t->lqua || !t->msymdd ||
      !t->msymddi || !t->msymdq || !t->msymdqi || !t->msymqq || !t->msymqqi || !t->rot_int || !t->rot_atom_abs) {
    set_error("fdmgpu_tensor_create: invalid argument (lseuil 0..3, 1 or 2 edges, lmax 0..3, a multipole, tables)");
    return FDMGPU_EINVAL;
  }
  const int nrang = (t->e1e2 || t->e2e2) ? 2 : 1;
  const int nip = t->rof_ip_max - t->rof_ip0 + 1;
  if (t->rof_ip0 > 1 || t->rof_ip_max < nrang || t->rof_lmax < t->lmax || t->rof_nbseuil < t->nbseuil ||
      t->taull_ld < (t->lmax + 1) * (t->lmax + 1) || t->taull_offset < 0) {
    set_error("fdmgpu_tensor_create: the radial integrals must cover ip 1.." + std::to_string(nrang) +
              ", L 0..lmax and every edge, and Taull the probe's harmonics");
    return FDMGPU_EINVAL;
  }
  int ninit1 = 0;
  const std::vector<CoreState> core = core_states(t->jseuil, t->lseuil, t->nbseuil, ninit1);
  const int ninit = (int)core.size();
  ninitr = t->core_resolved ? ninit : t->nbseuil;
  if (ninitr > TENS_MAXINIT) {
    set_error("fdmgpu_tensor_create: more than 16 core states resolved");
    return FDMGPU_EINVAL;
  }
  const int li = t->lseuil, lmax = t->lmax;
  nlm = (lmax + 1) * (lmax + 1);
  static const GauntTable G;
  double clm[3][3][3][16];
  clm_table(clm);

  std::map<std::tuple<int, int, int, int>, int> comp_of;
  std::vector<std::vector<std::vector<TensTerm>>> comp_terms;  // [comp][initr] terms in tens_ab's order
  auto component = [&](int irang, int jrang, int l_s, int m_s, int l_i, int m_i) {
    const auto key = std::make_tuple(irang, jrang, l_s * l_s + l_s + m_s, l_i * l_i + l_i + m_i);
    auto it = comp_of.find(key);
    if (it != comp_of.end()) return it->second;
    const int c = (int)comp_terms.size();
    comp_of[key] = c;
    comp_terms.emplace_back(ninitr);
    // tens_ab's loops (tensor.f90:989-1250) for this branch
    for (int ig = 0; ig < ninit; ig++) {
      const int iseuil = ig < ninit1 ? 0 : 1;
      const int initr = t->core_resolved ? ig : iseuil;
      for (int s1 = 0; s1 < 2; s1++) {
        const double c1 = core[ig].c[s1];
        if (std::fabs(c1) < 1e-6) continue;
        for (int s2 = 0; s2 < 2; s2++) {
          if (li == 0 && s1 != s2) continue;
          const double c2 = core[ig].c[s2];
          if (std::fabs(c2) < 1e-6) continue;
          if (s2 != s1) continue;  // one final-state spin per core spin (no spin-orbit, no spin flip)
          const double ci2 = c1 * c2;
          for (int lf1 = 0; lf1 <= lmax; lf1++) {
            if (t->lplus1 && lf1 < li + 1) continue;
            if (t->lmoins1 && lf1 > li) continue;
            if (lf1 > li + irang || lf1 < li - irang || lf1 % 2 != (li + irang) % 2) continue;
            for (int mf1 = -lf1; mf1 <= lf1; mf1++) {
              const cd gs = G.xrc(lf1, mf1, l_s, m_s, li, core[ig].m[s1]);
              if (std::abs(gs) < 1e-10) continue;
              for (int lf2 = 0; lf2 <= lmax; lf2++) {
                if (t->lplus1 && lf2 < li + 1) continue;
                if (t->lmoins1 && lf2 > li) continue;
                if (lf2 > li + jrang || lf2 < li - jrang || lf2 % 2 != (li + jrang) % 2) continue;
                for (int mf2 = -lf2; mf2 <= lf2; mf2++) {
                  const cd gi = G.xrc(lf2, mf2, l_i, m_i, li, core[ig].m[s2]);
                  if (std::abs(gi) < 1e-10) continue;
                  const cd cg = (ci2 * std::conj(gs)) * gi;
                  TensTerm q;
                  q.lm1 = lf1 * lf1 + lf1 + mf1;
                  q.lm2 = lf2 * lf2 + lf2 + mf2;
                  q.r1 = (iseuil * nip + irang - t->rof_ip0) * (t->rof_lmax + 1) + lf1;
                  q.r2 = (iseuil * nip + jrang - t->rof_ip0) * (t->rof_lmax + 1) + lf2;
                  q.cg_re = cg.real();
                  q.cg_im = cg.imag();
                  comp_terms[c][initr].push_back(q);
                }
              }
            }
          }
        }
      }
    }
    return c;
  };

  std::vector<int> rep_ptr{0}, rep_comp;
  std::vector<double> rep_w;
  std::vector<int> fill(4 * TENS_NOUT, 0);
  for (int o = 0; o < TENS_NOUT; o++) fill[4 * o] = -1;
  std::vector<char> filled(TENS_NOUT, 0);
  auto put = [&](int o, int rep, int is, int isi, int cj) {
    fill[4 * o] = rep;
    fill[4 * o + 1] = is;
    fill[4 * o + 2] = isi;
    fill[4 * o + 3] = cj;
    filled[o] = (is != 0 || isi != 0);
  };
  // tenseur_car's loops over the independent elements (tensor.f90:352-500)
  for (int irang = 1; irang <= nrang; irang++)
    for (int jrang = irang; jrang <= nrang; jrang++) {
      if (!t->e1e1 && irang == 1 && jrang == 1) continue;
      if (!t->e1e2 && irang == 1 && jrang == 2) continue;
      if (!t->e2e2 && irang == 2 && jrang == 2) continue;
      for (int ks = 0; ks < 3; ks++) {
        if (irang == 1 && t->ldip[ks] != 1) continue;
        for (int js = 0; js < (irang > 1 ? 3 : 1); js++) {
          if (irang == 2 && t->lqua[i3(ks, js)] != 1) continue;
          for (int ki = 0; ki < 3; ki++) {
            if (jrang == 1 && t->ldip[ki] != 1) continue;
            for (int ji = 0; ji < (jrang > 1 ? 3 : 1); ji++) {
              if (jrang == 2 && t->lqua[i3(ki, ji)] != 1) continue;
              int a0, b0;
              if (irang == 1 && jrang == 1) {
                if (ki < ks) continue;
                a0 = t->msymdd[i3(ks, ki)];
                b0 = t->msymddi[i3(ks, ki)];
                if ((a0 == 0 && b0 == 0) || filled[i3(ks, ki)]) continue;
              } else if (irang == 1) {
                a0 = t->msymdq[i4(ks, ki, ji)];
                b0 = t->msymdqi[i4(ks, ki, ji)];
                if ((a0 == 0 && b0 == 0) || filled[9 + i4(ks, ki, ji)]) continue;
              } else {
                a0 = t->msymqq[i5(ks, js, ki, ji)];
                b0 = t->msymqqi[i5(ks, js, ki, ji)];
                if ((a0 == 0 && b0 == 0) || filled[36 + i5(ks, js, ki, ji)]) continue;
              }
              const int rep = (int)rep_ptr.size() - 1;
              int lms = 0;
              for (int l_s = 0; l_s <= irang; l_s++)
                for (int m_s = -l_s; m_s <= l_s; m_s++, lms++) {
                  const double cs = clm[irang][ks][js][lms];
                  if (std::fabs(cs) < 1e-10) continue;
                  int lmi = 0;
                  for (int l_i = 0; l_i <= jrang; l_i++)
                    for (int m_i = -l_i; m_i <= l_i; m_i++, lmi++) {
                      const double ci = clm[jrang][ki][ji][lmi];
                      if (std::fabs(ci) < 1e-10) continue;
                      rep_comp.push_back(component(irang, jrang, l_s, m_s, l_i, m_i));
                      rep_w.push_back(cs * ci);
                    }
                }
              rep_ptr.push_back((int)rep_comp.size());
              auto sgn = [](int a, int a0_) { return a != 0 ? a / a0_ : 0; };
              if (irang == 1 && jrang == 1) {
                for (int kks = 0; kks < 3; kks++)
                  for (int kki = kks; kki < 3; kki++) {
                    const int a = t->msymdd[i3(kks, kki)], b = t->msymddi[i3(kks, kki)];
                    if (std::abs(a) != std::abs(a0) || std::abs(b) != std::abs(b0)) continue;
                    put(i3(kks, kki), rep, sgn(a, a0), sgn(b, b0), 0);
                    put(i3(kki, kks), rep, sgn(a, a0), sgn(b, b0), 1);
                  }
              } else if (irang == 1) {
                for (int kks = 0; kks < 3; kks++)
                  for (int kki = 0; kki < 3; kki++)
                    for (int jji = kki; jji < 3; jji++) {
                      const int a = t->msymdq[i4(kks, kki, jji)], b = t->msymdqi[i4(kks, kki, jji)];
                      if (std::abs(a) != std::abs(a0) || std::abs(b) != std::abs(b0)) continue;
                      put(9 + i4(kks, kki, jji), rep, sgn(a, a0), sgn(b, b0), 0);
                      put(9 + i4(kks, jji, kki), rep, sgn(a, a0), sgn(b, b0), 0);
                    }
              } else {
                for (int kks = 0; kks < 3; kks++)
                  for (int jjs = 0; jjs < 3; jjs++)
                    for (int kki = 0; kki < 3; kki++)
                      for (int jji = 0; jji < 3; jji++) {
                        if (filled[36 + i5(kks, jjs, kki, jji)]) continue;
                        const int a = t->msymqq[i5(kks, jjs, kki, jji)], b = t->msymqqi[i5(kks, jjs, kki, jji)];
                        if (std::abs(a) != std::abs(a0) || std::abs(b) != std::abs(b0)) continue;
                        const int is = sgn(a, a0), isi = sgn(b, b0);
                        put(36 + i5(kks, jjs, kki, jji), rep, is, isi, 0);
                        put(36 + i5(jjs, kks, kki, jji), rep, is, isi, 0);
                        put(36 + i5(kks, jjs, jji, kki), rep, is, isi, 0);
                        put(36 + i5(jjs, kks, jji, kki), rep, is, isi, 0);
                        put(36 + i5(kki, jji, kks, jjs), rep, is, isi, 1);
                        put(36 + i5(jji, kki, kks, jjs), rep, is, isi, 1);
                        put(36 + i5(kki, jji, jjs, kks), rep, is, isi, 1);
                        put(36 + i5(jji, kki, jjs, kks), rep, is, isi, 1);
                      }
              }
            }
          }
        }
      }
    }
  const int ncomp = (int)comp_terms.size(), nrep = (int)rep_ptr.size() - 1;
  std::vector<int> term_ptr{0};
  std::vector<TensTerm> terms;
  for (int c = 0; c < ncomp; c++)
    for (int k = 0; k < ninitr; k++) {
      terms.insert(terms.end(), comp_terms[c][k].begin(), comp_terms[c][k].end());
      term_ptr.push_back((int)terms.size());
    }
  const size_t lds = sizeof(z2) * ((size_t)std::max(ncomp, TENS_NOUT) + std::max(nrep, 1)) * ninitr;
  if (lds > 65536) {
    set_error("fdmgpu_tensor_create: the tensors of this edge need more than 64 KiB of LDS");
    return FDMGPU_EINVAL;
  }
  // rot_int x rot_atom_abs^T
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double s0 = 0;
      for (int k = 0; k < 3; k++) s0 += t->rot_int[i3(i, k)] * t->rot_atom_abs[i3(j, k)];
      d.rot[i3(i, j)] = s0;
    }
  int *dtp = nullptr, *drp = nullptr, *drc = nullptr, *dfill = nullptr;
  TensTerm *dterms = nullptr;
  double *drw = nullptr;
  if (to_device(term_ptr, &dtp, owned) || to_device(terms, &dterms, owned) || to_device(rep_ptr, &drp, owned) ||
      to_device(rep_comp, &drc, owned) || to_device(rep_w, &drw, owned) || to_device(fill, &dfill, owned))
    return FDMGPU_EHIP;
  d.ncomp = ncomp;
  d.nrep = nrep;
  d.ninitr = ninitr;
  d.hydrogen = t->numat == 1;
  d.nrof = t->rof_nbseuil * nip * (t->rof_lmax + 1);
  d.term_ptr = dtp;
  d.terms = dterms;
  d.rep_ptr = drp;
  d.rep_comp = drc;
  d.rep_w = drw;
  d.fill = dfill;
  d.out_dd = this->dd = t->e1e1 != 0;
  d.out_dq = this->dq = t->e1e2 != 0;
  d.out_qq = this->qq = t->e2e2 != 0;
  nrof = d.nrof;
  toff = t->taull_offset;
  ld = t->taull_ld;
  max_points = maxp;
  stream = s;
  return 0;
}

int TensorPlan::run(int count, const TensPoints &P) {
  if (count < 1 || count > max_points) {
    set_error("fdmgpu_tensor_run: count must be in 1.." + std::to_string(max_points));
    return FDMGPU_EINVAL;
  }
  for (int k = 0; k < count; k++)
    if (!P.rof[k] || !P.taull[k] || (dd && !P.dd[k]) || (dq && !P.dq[k]) || (qq && !P.qq[k])) {
      set_error("fdmgpu_tensor_run: null buffer for point " + std::to_string(k));
      return FDMGPU_EINVAL;
    }
  const size_t lds = sizeof(z2) * ((size_t)std::max(d.ncomp, TENS_NOUT) + std::max(d.nrep, 1)) * ninitr;
  hipLaunchKernelGGL(k_tensor, dim3(count), dim3(256), lds, stream, d, P, toff, ld);
  FDM_HIP(hipGetLastError());
  return 0;
}

}  // namespace fdmgpu
