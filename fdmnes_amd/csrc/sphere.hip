// The atoms' radial solutions on the device (Sphere, sphere.f90:8-210): for
// every energy point of a call, every atom and every L, the radial
// Schrodinger equation integrated outwards on the atom's radial grid
// (coef_sch_rad :324, Sch_radial :413: series start at the first two
// points, three-point recurrence), matched to the normalised Bessel and
// Hankel functions at the muffin-tin radius (Renormal :931, cal_ampl
// :1157), then phiato at the atom's border points (cal_phiato :3186:
// quadratic interpolation at the point's distance, times the real harmonic
// there).  One spin, spherical potential, no spin-orbit or Hubbard term,
// the regular solution.  Multiply-adds are not fused (contraction off) and
// every real operation follows the reference's order; the complex square
// root, exponential, sine, cosine and divisions are this file's own.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "sphere.h"
#include "solver.h"

#pragma clang fp contract(off)

namespace fdmgpu {
namespace {

struct cz {
  double re, im;
};
__device__ inline cz cadd(cz a, cz b) { return {a.re + b.re, a.im + b.im}; }
__device__ inline cz csub(cz a, cz b) { return {a.re - b.re, a.im - b.im}; }
__device__ inline cz cmul(cz a, cz b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ inline cz cscale(cz a, double s) { return {a.re * s, a.im * s}; }
// a / b by Smith's algorithm (R. L. Smith, CACM 5(8):435, 1962)
__device__ inline cz cdiv(cz a, cz b) {
  if (fabs(b.re) < fabs(b.im)) {
    const double r = b.re / b.im, d = b.re * r + b.im;
    return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
  }
  const double r = b.im / b.re, d = b.im * r + b.re;
  return {(a.im * r + a.re) / d, (a.im - a.re * r) / d};
}
__device__ inline cz csqrt_(cz z) {  // principal square root
  if (z.re == 0 && z.im == 0) return {0, z.im};
  const double t = sqrt((fabs(z.re) + hypot(z.re, z.im)) * 0.5);
  if (z.re >= 0) return {t, z.im / (2 * t)};
  return {fabs(z.im) / (2 * t), copysign(t, z.im)};
}
__device__ inline cz cexp_(cz z) {
  const double e = exp(z.re);
  return {e * cos(z.im), e * sin(z.im)};
}
__device__ inline cz csin_(cz z) { return {sin(z.re) * cosh(z.im), cos(z.re) * sinh(z.im)}; }
__device__ inline cz ccos_(cz z) { return {cos(z.re) * cosh(z.im), -(sin(z.re) * sinh(z.im))}; }

constexpr double kPi = 3.1415926535897932384626433832;
constexpr double kA2s4 = 0.25 * (0.0072973525664 * 0.0072973525664);
constexpr double kAlfa = 0.0072973525664;

// One thread per (point, atom, L): the integration, the matching and the
// renormalised solution written to u (re, im per radial point)
__global__ void k_sph_radial(SphTabs T, SphPoints P, int lmax_all, double *__restrict__ u, int64_t ustride,
                             double *__restrict__ uirr) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x, g = blockIdx.y;
  if (t >= T.natome * (lmax_all + 1)) return;
  const int a = t / (lmax_all + 1), L = t % (lmax_all + 1);
  if (L > T.lmax[a]) return;
  const int n = T.nr[a];
  const double *r = T.r + T.roff[a], *V = T.V + T.roff[a];
  const double *clm = T.clm + T.roff[a], *clp = T.clp + T.roff[a], *cl0 = T.cl0 + T.roff[a];
  const double *cgm = T.cgm + T.roff[a], *cgp = T.cgp + T.roff[a], *cg0 = T.cg0 + T.roff[a];
  const double *f2 = T.f2 + T.roff[a];
  double *ur = u + (((int64_t)g * T.natome + a) * (lmax_all + 1) + L) * ustride;
  double *ui = ur + T.nrmax;
  const double Ec = P.ecin[g], Ei = P.eimag[g], Ev = P.ener[g];
  const int Z = T.Z[a];
  const bool ecomp = fabs(Ei) > 1e-10 || Ec < 1e-10;
  // coefficients of row i (coef_sch_rad, sphere.f90:371-400); gp as its inverse
  auto coef = [&](int i, double &g0, double &gm, double &gpi) {
    const double vme = V[i] - Ev;
    g0 = -cl0[i] + vme;
    gm = -clm[i];
    double gp = -clp[i];
    if (T.rel) {
      const double bder = 1 / (1 - kA2s4 * vme);
      double dvr;
      if (i == 0)
        dvr = (double)(2 * Z) / (r[0] * r[0]);
      else if (i == n - 1)
        dvr = cgm[i] * V[i - 1] + cg0[i] * V[i] + cgp[i] * V[i];
      else
        dvr = cgm[i] * V[i - 1] + cg0[i] * V[i] + cgp[i] * V[i + 1];
      const double fac = kA2s4 * bder * dvr;
      g0 = g0 - kA2s4 * (vme * vme) - fac * (cg0[i] - 1 / r[i]);
      gm = gm - fac * cgm[i];
      gp = gp - fac * cgp[i];
    }
    gpi = 1 / gp;
  };
  // series at the origin (Sch_radial, sphere.f90:461-551)
  const double ci = -Ei / (4 * L + 6);
  const double br = -(Z / (L + 1.0));
  const double cr = -((double)(2 * Z) * br + Ec) / (4 * L + 6);
  const double p = T.rel ? sqrt((double)(L * L + L + 1) - (kAlfa * Z) * (kAlfa * Z)) : L + 1.0;
  for (int k = 0; k < 2; k++) {
    ur[k] = 1.0 * (1 + br * r[k] + cr * (r[k] * r[k])) * pow(r[k], p);
    ui[k] = ecomp ? 1.0 * ci * pow(r[k], p + 2) : 0.0;
  }
  // outward three-point recurrence (:576-700)
  const int L2 = L * (L + 1);
  double um = ur[0], u0 = ur[1], vm = ui[0], v0 = ui[1];
  for (int ir = 1; ir <= n - 2; ir++) {
    double g0, gm, gpi;
    coef(ir, g0, gm, gpi);
    const double td = g0 + L2 * f2[ir];
    double x = td * u0 + gm * um;
    double y = 0;
    if (ecomp) {
      y = td * v0 + gm * vm;
      x = x + Ei * v0;
      y = y - Ei * u0;
    }
    const double up = -x * gpi, vp = ecomp ? -y * gpi : 0.0;
    ur[ir + 1] = up;
    ui[ir + 1] = vp;
    um = u0;
    u0 = up;
    vm = v0;
    v0 = vp;
  }
  for (int i = 0; i < n; i++) {
    ur[i] = ur[i] / r[i];
    if (ecomp) ui[i] = ui[i] / r[i];
  }
  // matching at rmtg (Renormal, :954-1061; cal_ampl :1177-1188)
  const double rm = T.rmtg[a];
  const int inr = T.inr[a];  // 1-based min(nrmtg, nr)
  const double d12 = r[inr - 2] - r[inr - 3], d01 = r[inr - 1] - r[inr - 2], d02 = r[inr - 1] - r[inr - 3];
  const double d2 = 1 / (d12 * d02), d1 = 1 / (d12 * d01), d0 = 1 / (d01 * d02);
  const double A = ur[inr - 3] * d2 - ur[inr - 2] * d1 + ur[inr - 1] * d0;
  const double B = (ur[inr - 2] - ur[inr - 3]) / d12 - A * (r[inr - 2] + r[inr - 3]);
  const double Cc = ur[inr - 1] - A * (r[inr - 1] * r[inr - 1]) - B * r[inr - 1];
  double Ai = 0, Bi = 0, Ci = 0;
  if (ecomp) {
    Ai = ui[inr - 3] * d2 - ui[inr - 2] * d1 + ui[inr - 1] * d0;
    Bi = (ui[inr - 2] - ui[inr - 3]) / d12 - Ai * (r[inr - 2] + r[inr - 3]);
    Ci = ui[inr - 1] - Ai * (r[inr - 1] * r[inr - 1]) - Bi * r[inr - 1];
  }
  const cz u1 = {A * (rm * rm) + B * rm + Cc, Ai * (rm * rm) + Bi * rm + Ci};
  const cz u2 = {2 * A * rm + B, 2 * Ai * rm + Bi};
  // normalised Hankel and Bessel functions of order L and L + 1 at rmtg
  // (Cal_Hankel, Cal_Bessel, sphere.f90:7255-7350)
  const cz k = csqrt_({Ec, Ei});
  const cz z = cscale(k, rm);
  cz h0 = {0, -1}, h1 = csub({-1, 0}, cdiv({0, 1}, z));
  cz b0 = csin_(z), b1 = cadd(cscale(ccos_(z), -1.0), cdiv(b0, z));
  for (int l = 2; l <= L + 1; l++) {  // h_l of l = L and L + 1 kept in h0, h1
    const cz hn = csub(cdiv(cscale(h1, (double)(2 * l - 1)), z), h0);
    const cz bn = csub(cdiv(cscale(b1, (double)(2 * l - 1)), z), b0);
    h0 = h1;
    h1 = hn;
    b0 = b1;
    b1 = bn;
  }
  const cz fnorm = csqrt_({k.re / kPi, k.im / kPi});
  const cz cfh = cdiv(cmul(fnorm, cexp_({-z.im, z.re})), z), cfb = cdiv(fnorm, z);
  const cz dh = csub(cdiv(cscale(h0, (double)L), z), h1), db = csub(cdiv(cscale(b0, (double)L), z), b1);
  const cz hs = cmul(cfh, h0), dhs = cmul(cmul(cfh, k), dh);
  const cz bs = cmul(cfb, b0), dbs = cmul(cmul(cfb, k), db);
  const double wout = 1 / (kPi * (rm * rm));
  const cz s1 = {hs.im, -hs.re}, s2 = {dhs.im, -dhs.re};  // -i h, -i h'
  const cz ws = csub(cmul(u1, s2), cmul(u2, s1)), we = csub(cmul(u1, dbs), cmul(u2, bs));
  const cz ampl = cdiv({wout, 0.0}, we);
  if (P.tau[g]) {
    const cz tau = cdiv(cscale(we, -1.0), ws);
    double *o = P.tau[g] + 2 * (T.toff[a] + L);
    o[0] = tau.re;
    o[1] = tau.im;
  }
  for (int i = 0; i < n; i++) {
    const cz w = cmul(ampl, {ur[i], ecomp ? ui[i] : 0.0});
    ur[i] = w.re;
    ui[i] = ecomp ? w.im : 0.0;
  }
  if (!P.phi2[g]) return;
  // the irregular solution (Sch_radial_solsing, sphere.f90:2591-2806, as
  // Sphere calls it at :152-162): -i r h_L(k r) from the matching point
  // outwards, h = fnorm (j + i n) by cbessneu's upward recurrence
  // (:7226-7248; its real form :7356-7378 for real energies), then inwards
  // by the same three-point recurrence with gm and gp exchanged; over r
  double *sr = uirr + (((int64_t)g * T.natome + a) * (lmax_all + 1) + L) * ustride;
  double *si = sr + T.nrmax;
  const cz fnc = csqrt_({k.re / kPi, k.im / kPi});
  const double kr = k.re, fnr = sqrt(kr / kPi);
  for (int ir = inr; ir <= n; ir++) {
    const double x = r[ir - 1];
    cz bl, nl;
    if (ecomp) {
      const cz zz = cscale(k, x);
      cz n0 = cscale(cdiv(cmul(fnc, ccos_(zz)), zz), -1.0), b0 = cdiv(cmul(fnc, csin_(zz)), zz);
      bl = b0;
      nl = n0;
      if (L >= 1) {
        cz n1 = csub(cdiv(n0, zz), b0), b1 = cadd(cdiv(b0, zz), n0);
        for (int l = 2; l <= L; l++) {
          const cz n2 = csub(cdiv(cscale(n1, (double)(2 * l - 1)), zz), n0);
          const cz b2 = csub(cdiv(cscale(b1, (double)(2 * l - 1)), zz), b0);
          n0 = n1;
          n1 = n2;
          b0 = b1;
          b1 = b2;
        }
        bl = b1;
        nl = n1;
      }
    } else {
      const double zz = kr * x;
      double n0 = -fnr * cos(zz) / zz, b0 = fnr * sin(zz) / zz;
      bl = {b0, 0.0};
      nl = {n0, 0.0};
      if (L >= 1) {
        double n1 = n0 / zz - b0, b1 = b0 / zz + n0;
        for (int l = 2; l <= L; l++) {
          const double n2 = (2 * l - 1) * n1 / zz - n0, b2 = (2 * l - 1) * b1 / zz - b0;
          n0 = n1;
          n1 = n2;
          b0 = b1;
          b1 = b2;
        }
        bl = {b1, 0.0};
        nl = {n1, 0.0};
      }
    }
    const cz h = {bl.re - nl.im, bl.im + nl.re};  // bess + i neum
    sr[ir - 1] = x * h.im;                        // -i r h
    si[ir - 1] = -(x * h.re);
  }
  for (int ir = inr; ir >= 2; ir--) {
    const int i0 = ir - 1, ip = ir - 2, im = ir;  // 0-based: the point, the next one inwards, the one outwards
    const double vme = V[i0] - Ev;
    double g0 = -cl0[i0] + vme, gm = -clm[i0], gp = -clp[i0];
    if (T.rel) {
      const double bder = 1 / (1 - kA2s4 * vme);
      double dvr;
      if (i0 == 0)
        dvr = (double)(2 * Z) / (r[0] * r[0]);
      else if (i0 == n - 1)
        dvr = cgm[i0] * V[i0 - 1] + cg0[i0] * V[i0] + cgp[i0] * V[i0];
      else
        dvr = cgm[i0] * V[i0 - 1] + cg0[i0] * V[i0] + cgp[i0] * V[i0 + 1];
      const double fac = kA2s4 * bder * dvr;
      g0 = g0 - kA2s4 * (vme * vme) - fac * (cg0[i0] - 1 / r[i0]);
      gm = gm - fac * cgm[i0];
      gp = gp - fac * cgp[i0];
    }
    const double td = g0 + L2 * f2[i0], gmi = 1 / gm;
    double x = td * sr[i0] + gp * sr[im], y = td * si[i0] + gp * si[im];
    if (ecomp) {  // - i Eimag us(ir)
      x = x + Ei * si[i0];
      y = y - Ei * sr[i0];
    }
    sr[ip] = -x * gmi;
    si[ip] = -y * gmi;
  }
  for (int i = 0; i < n; i++) {
    sr[i] = sr[i] / r[i];
    si[i] = si[i] / r[i];
  }
}

// One thread per (point, border point of all atoms, Lm): phiato by
// quadratic interpolation of the renormalised solution of l(Lm) at the
// point's distance, times its real harmonic (cal_phiato, :3224-3345)
__global__ void k_sph_phiato(SphTabs T, SphPoints P, int lmax_all, int nlm_max, const double *__restrict__ u,
                             int64_t ustride, int cplx, const double *__restrict__ u2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x, g = blockIdx.y;
  if (t >= T.nball * nlm_max) return;
  const int q = t / nlm_max, Lm = t % nlm_max;
  const int a = T.batom[q];
  const int nl = (T.lmax[a] + 1) * (T.lmax[a] + 1);
  if (Lm >= nl) return;
  int L = 0;
  while ((L + 1) * (L + 1) <= Lm) L++;
  const double *r = T.r + T.roff[a];
  const double *ur = u + (((int64_t)g * T.natome + a) * (lmax_all + 1) + L) * ustride;
  const double *ui = ur + T.nrmax;
  const int i = T.bint[q];  // 1-based upper point of the three
  const double rr = T.rrel[q], rmm = r[i - 3], r0 = r[i - 2], rp = r[i - 1];
  auto interp = [&](double fm, double f0, double fp) {  // f_interp2, tools.f90:504-522
    const double dp0 = rp - r0, dpm = rp - rmm, d0m = r0 - rmm;
    const double A = (fm * dp0 - f0 * dpm + fp * d0m) / (d0m * dp0 * dpm);
    const double B = (f0 - fm) / d0m - A * (r0 + rmm);
    const double C = f0 - A * (r0 * r0) - B * r0;
    return A * (rr * rr) + B * rr + C;
  };
  const double ph = interp(ur[i - 3], ur[i - 2], ur[i - 1]);
  const double phi = interp(ui[i - 3], ui[i - 2], ui[i - 1]);
  const double y = T.ylmato[T.yoff[q] + Lm];
  const int64_t o = T.yoff[q] + Lm;  // the same (border point, Lm) place in phiato
  if (cplx) {
    P.phi[g][2 * o] = ph * y;
    P.phi[g][2 * o + 1] = phi * y;
  } else {
    P.phi[g][o] = ph * y;
  }
  if (P.phi2[g]) {  // the irregular solution (i_RI = 2); its imaginary part only for complex radial solutions
    const bool ecomp = fabs(P.eimag[g]) > 1e-10 || P.ecin[g] < 1e-10;
    const double *sr = u2 + (((int64_t)g * T.natome + a) * (lmax_all + 1) + L) * ustride, *si = sr + T.nrmax;
    const double p2 = interp(sr[i - 3], sr[i - 2], sr[i - 1]);
    const double p2i = ecomp ? interp(si[i - 3], si[i - 2], si[i - 1]) : 0.0;
    if (cplx) {
      P.phi2[g][2 * o] = p2 * y;
      P.phi2[g][2 * o + 1] = p2i * y;
    } else {
      P.phi2[g][o] = p2 * y;
    }
  }
}

// The absorbing atom's radial integrals of the golden rule (Radial_integral
// sphere.f90:1490 -> radial_matrix :2209): one wave per (point, edge, ip, L),
// rof = fac(ip) * integral_0^rmtsd psii r u_L r^ip dr by f_integr3's
// piecewise cubics (tools.f90:308-419), every interval's share computed
// independently from the four points its cubic goes through, the shares
// summed across the wave.
__device__ inline double integr3_share(const double *r, int i, int nlast, int nrm, double radius, double f1,
                                       double f2, double f3, double f4, double f5) {
  // i: 1-based interval; f1..f5 the integrand at points i-2 .. i+2 (those outside the grid unused)
  double r1, r2, r3, r4, g1, g2, g3, g4, rm, rp;
  if (i == 1) {  // from 0, with the value at 0 from the parabola through points 1..3
    if (fabs(f5) < 1e-20 && fabs(f4) < 1e-20 && fabs(f3) < 1e-20) return 0.0;
    const double rmm = r[0], r0 = r[1], rpp = r[2];
    const double dp0 = rpp - r0, dpm = rpp - rmm, d0m = r0 - rmm;
    const double a = (f3 * dp0 - f4 * dpm + f5 * d0m) / (d0m * dp0 * dpm);
    const double b = (f4 - f3) / d0m - a * (r0 + rmm);
    g1 = f4 - a * (r0 * r0) - b * r0;
    r1 = 0;
    r2 = r[0]; g2 = f3;
    r3 = r[1]; g3 = f4;
    r4 = r[2]; g4 = f5;
    rm = 0;
    rp = r3;
  } else if (i == nlast) {  // cut at radius (or the grid's end), with the cubic of interval i - 1
    rm = r[i - 1];
    rp = r[nrm - 1];
    if (i < 3) {
      r1 = r[i - 2]; g1 = f2;
      r2 = r[i - 1]; g2 = f3;
      r3 = r[i]; g3 = f4;
      r4 = r[i + 1]; g4 = f5;
    } else {
      r1 = r[i - 3]; g1 = f1;
      r2 = r[i - 2]; g2 = f2;
      r3 = r[i - 1]; g3 = f3;
      r4 = r[i]; g4 = f4;
    }
  } else {
    r1 = r[i - 2]; g1 = f2;
    r2 = r[i - 1]; g2 = f3;
    r3 = r[i]; g3 = f4;
    r4 = r[i + 1]; g4 = f5;
    rm = r2;
    rp = r3;
  }
  if (rp > radius) rp = radius;
  if (fabs(g2) < 1e-20 && fabs(g3) < 1e-20) return 0.0;
  const double rap14 = (g1 - g4) / (r1 - r4), rap24 = (g2 - g4) / (r2 - r4), rap34 = (g3 - g4) / (r3 - r4);
  const double fac1 = (rap14 - rap34) / (r1 - r3), fac2 = (rap24 - rap34) / (r2 - r3);
  const double a = (fac1 - fac2) / (r1 - r2);
  const double b = fac1 - a * (r1 + r3 + r4);
  const double c = rap14 - a * (r1 * r1 + r1 * r4 + r4 * r4) - b * (r1 + r4);
  const double d = g1 - (a * (r1 * r1) + b * r1 + c) * r1;
  return (0.25 * (rm + rp) * (rm * rm + rp * rp) * a + (1.0 / 3) * (rm * rm + rm * rp + rp * rp) * b +
          0.5 * (rm + rp) * c + d) *
         (rp - rm);
}

__global__ void __launch_bounds__(64) k_sph_rof(SphTabs T, SphPoints P, RofTabs R, int lmax_all,
                                                const double *__restrict__ u, int64_t ustride) {
  const int g = blockIdx.y, nip = R.ip_max - R.ip0 + 1;
  int t = blockIdx.x;
  const int L = t % (R.lmax + 1);
  t /= R.lmax + 1;
  const int ipo = t % nip, is = t / nip, ip = R.ip0 + ipo;
  const int a = R.atom, n = T.nr[a];
  const double *r = T.r + T.roff[a];
  const double *ur = u + (((int64_t)g * T.natome + a) * (lmax_all + 1) + L) * ustride;
  const double *ui = ur + T.nrmax;
  const double *ps = R.psii + (int64_t)is * n;
  const bool ecomp = fabs(P.eimag[g]) > 1e-10 || P.ecin[g] < 1e-10;
  // the integrand at 1-based point j (0 outside the grid; such values are never used)
  auto F = [&](const double *uu, int j) {
    if (j < 1 || j > n) return 0.0;
    const double x = r[j - 1];
    const double rb = ip == 0 ? 1.0 : ip == 1 ? x : ip == 2 ? x * x : x * x * x;
    return ps[j - 1] * x * uu[j - 1] * rb;
  };
  double sr = 0, si = 0;
  for (int i = 1 + (int)threadIdx.x; i <= R.nlast; i += 64) {
    sr += integr3_share(r, i, R.nlast, n, R.rmtsd, F(ur, i - 2), F(ur, i - 1), F(ur, i), F(ur, i + 1),
                        F(ur, i + 2));
    if (ecomp)
      si += integr3_share(r, i, R.nlast, n, R.rmtsd, F(ui, i - 2), F(ui, i - 1), F(ui, i), F(ui, i + 1),
                          F(ui, i + 2));
  }
  for (int off = 32; off > 0; off >>= 1) {
    sr += __shfl_down(sr, off);
    si += __shfl_down(si, off);
  }
  if (threadIdx.x == 0) {
    const double vecond = 0.5 * kAlfa * fmax(0.001, P.energ[g] + R.eseuil[is]);
    const double fac = ip == 0 ? 0.5 * kAlfa : ip == 1 ? 1.0 : ip == 2 ? 0.5 * vecond : -(1.0 / 6) * (vecond * vecond);
    double *o = P.rof[g] + 2 * (((int64_t)is * nip + ipo) * (R.lmax + 1) + L);
    o[0] = fac * sr;
    o[1] = fac * si;
  }
}

template <class T>
int upload(T **d, const T *h, int64_t n) {
  FDM_HIP(hipMalloc(d, sizeof(T) * (n > 0 ? n : 1)));
  if (n > 0) FDM_HIP(hipMemcpy(*d, h, sizeof(T) * n, hipMemcpyHostToDevice));
  return 0;
}

}  // namespace

SphereSolver::~SphereSolver() {
  for (void *p : {(void *)d.r, (void *)d.V, (void *)d.clm, (void *)d.clp, (void *)d.cl0, (void *)d.cgm, (void *)d.cgp,
                  (void *)d.cg0, (void *)d.f2, (void *)d.roff, (void *)d.nr, (void *)d.Z, (void *)d.lmax,
                  (void *)d.rmtg, (void *)d.inr, (void *)d.toff, (void *)d.batom, (void *)d.bint, (void *)d.rrel,
                  (void *)d.yoff, (void *)d.ylmato, (void *)u, (void *)d_psii, (void *)u2})
    if (p) (void)hipFree(p);
}

int SphereSolver::init(const fdmgpu_sphere_tables *t, int maxp, hipStream_t s) {
  if (t->natome < 1 || maxp < 1 || maxp > SPH_MAXP || !t->nr || !t->r || !t->V || !t->Z || !t->lmax || !t->rmtg ||
      !t->nbord || !t->bxyz || !t->pos || !t->ylmato) {
    set_error("fdmgpu_sphere_create: invalid argument");
    return FDMGPU_EINVAL;
  }
  const int na = t->natome;
  std::vector<int64_t> roff(na), toff(na), yoff;
  std::vector<int> inr(na), batom, bint;
  std::vector<double> clm, clp, cl0, cgm, cgp, cg0, f2, rr;
  int64_t nrt = 0, nt = 0, ny = 0;
  lmax_all = 0;
  nrmax = 0;
  for (int a = 0; a < na; a++) {
    const int n = t->nr[a];
    if (n < 4 || t->lmax[a] < 0 || t->lmax[a] > 12 || t->nbord[a] < 0 || !(t->rmtg[a] > 0)) {
      set_error("fdmgpu_sphere_create: atom " + std::to_string(a) + " needs nr >= 4, 0 <= lmax <= 12, rmtg > 0");
      return FDMGPU_EINVAL;
    }
    const double *r = t->r + nrt;
    for (int i = 0; i + 1 < n; i++)
      if (!(r[i] > 0 && r[i + 1] > r[i])) {
        set_error("fdmgpu_sphere_create: the radial grid of atom " + std::to_string(a) + " must increase from r > 0");
        return FDMGPU_EINVAL;
      }
    h_nr.push_back(n);
    h_lmax.push_back(t->lmax[a]);
    h_r.emplace_back(r, r + n);
    roff[a] = nrt;
    toff[a] = nt;
    nt += t->lmax[a] + 1;
    lmax_all = std::max(lmax_all, t->lmax[a]);
    nrmax = std::max(nrmax, n);
    // radial stencil (coef_sch_rad, sphere.f90:341-363)
    for (int i = 1; i <= n; i++) {
      const double rp = (i == n) ? r[n - 1] * r[n - 1] / r[n - 2] : r[i];
      const double rm = (i == 1) ? r[0] * r[0] / r[1] : r[i - 2];
      const double r0 = r[i - 1];
      const double dr = 0.5 * (rp - rm);
      const double m = 1 / ((r0 - rm) * dr), p = 1 / ((rp - r0) * dr);
      clm.push_back(m);
      clp.push_back(p);
      cl0.push_back(-m - p);
      const double gm = (rp - r0) / ((rm - r0) * (rp - rm)), gp = (rm - r0) / ((rp - r0) * (rm - rp));
      cgm.push_back(gm);
      cgp.push_back(gp);
      cg0.push_back(-gm - gp);
      f2.push_back(1 / (r[i - 1] * r[i - 1]));
    }
    // the matching point (mod_V, :243-246): the first r above rmtg + 1e-10
    int nrmtg;
    for (nrmtg = 1; nrmtg <= n - 1; nrmtg++)
      if (r[nrmtg - 1] > t->rmtg[a] + 1e-10) break;
    inr[a] = std::min(nrmtg, n);
    h_inr.push_back(inr[a]);
    if (inr[a] < 3) {
      set_error("fdmgpu_sphere_create: rmtg of atom " + std::to_string(a) + " below the third radial point");
      return FDMGPU_EINVAL;
    }
    // border points: distance (posrel), the interpolation's upper point (cal_phiato, :3214-3222)
    const int nl = (t->lmax[a] + 1) * (t->lmax[a] + 1);
    for (int b = 0; b < t->nbord[a]; b++) {
      const double *x = t->bxyz + 3 * (int64_t)batom.size();
      const double u0 = x[0] - t->pos[3 * a], u1 = x[1] - t->pos[3 * a + 1], u2 = x[2] - t->pos[3 * a + 2];
      const double d = std::sqrt(u0 * u0 + u1 * u1 + u2 * u2);
      int i;
      for (i = 2; i <= n; i++)
        if (r[i - 1] > d) break;
      i = std::max(std::min(i, n), 3);
      batom.push_back(a);
      bint.push_back(i);
      rr.push_back(d);
      yoff.push_back(ny);
      ny += nl;
    }
    nrt += n;
  }
  natome = na;
  nball = (int)batom.size();
  nphi = ny;
  ntau = nt;
  max_points = maxp;
  stream = s;
  d.natome = na;
  d.nball = nball;
  d.nrmax = nrmax;
  d.rel = t->relativiste != 0;
  std::vector<double> V(t->V, t->V + nrt);
  if (t->V_intmax < 1000.0)
    for (double &v : V) v = std::min(v, t->V_intmax);
  if (upload(&d.r, t->r, nrt) || upload(&d.V, (const double *)V.data(), nrt) || upload(&d.clm, clm.data(), nrt) ||
      upload(&d.clp, clp.data(), nrt) || upload(&d.cl0, cl0.data(), nrt) || upload(&d.cgm, cgm.data(), nrt) ||
      upload(&d.cgp, cgp.data(), nrt) || upload(&d.cg0, cg0.data(), nrt) || upload(&d.f2, f2.data(), nrt) ||
      upload(&d.roff, roff.data(), na) || upload(&d.nr, t->nr, na) || upload(&d.Z, t->Z, na) ||
      upload(&d.lmax, t->lmax, na) || upload(&d.rmtg, t->rmtg, na) || upload(&d.inr, inr.data(), na) ||
      upload(&d.toff, toff.data(), na) || upload(&d.batom, batom.data(), nball) ||
      upload(&d.bint, bint.data(), nball) || upload(&d.rrel, rr.data(), nball) || upload(&d.yoff, yoff.data(), nball) ||
      upload(&d.ylmato, t->ylmato, ny))
    return FDMGPU_EHIP;
  ustride = 2 * (int64_t)nrmax;
  FDM_HIP(hipMalloc(&u, sizeof(double) * (size_t)maxp * na * (lmax_all + 1) * ustride));
  return 0;
}

int SphereSolver::set_integrals(int atom, int nbs, const double *psii, const double *Eseuil, double rmtsd, int ip0,
                                int ip_max) {
  if (atom < 0 || atom >= natome || nbs < 1 || nbs > 2 || !psii || !Eseuil || ip0 < 0 || ip_max > 3 ||
      ip_max < ip0) {
    set_error("fdmgpu_sphere_set_integrals: atom in 0..natome-1, 1 or 2 edges, 0 <= ip0 <= ip_max <= 3");
    return FDMGPU_EINVAL;
  }
  const std::vector<double> &r = h_r[atom];
  const int n = h_nr[atom];
  if (!(rmtsd > r[2]) || !(rmtsd < 1e30)) {
    set_error("fdmgpu_sphere_set_integrals: rmtsd must lie beyond the atom's third radial point");
    return FDMGPU_EINVAL;
  }
  // f_integr3's last interval: the first i >= 2 with r(i+1) > rmtsd, or nr - 1
  int nlast = n - 1;
  for (int i = 2; i <= n - 1; i++)
    if (i == n - 1 || r[i] > rmtsd) {
      nlast = i;
      break;
    }
  if (d_psii) {
    FDM_HIP(hipFree(d_psii));
    d_psii = nullptr;
  }
  if (upload(&d_psii, psii, (int64_t)nbs * n)) return FDMGPU_EHIP;
  rt = RofTabs{};
  rt.atom = atom;
  rt.lmax = h_lmax[atom];
  rt.nbs = nbs;
  rt.ip0 = ip0;
  rt.ip_max = ip_max;
  rt.nlast = nlast;
  rt.eseuil[0] = Eseuil[0];
  rt.eseuil[1] = nbs > 1 ? Eseuil[1] : 0.0;
  rt.psii = d_psii;
  rt.rmtsd = rmtsd;
  return 0;
}

int SphereSolver::run(int count, const SphPoints &P, bool cplx) {
  if (count < 1 || count > max_points) {
    set_error("fdmgpu_sphere_run: count must be in 1.." + std::to_string(max_points));
    return FDMGPU_EINVAL;
  }
  bool want_rof = false;
  for (int k = 0; k < count; k++) want_rof = want_rof || P.rof[k] != nullptr;
  if (want_rof) {
    if (rt.atom < 0) {
      set_error("fdmgpu_sphere_run_integrals: no absorbing atom (fdmgpu_sphere_set_integrals)");
      return FDMGPU_EINVAL;
    }
    for (int k = 0; k < count; k++)
      if (!P.rof[k]) {
        set_error("fdmgpu_sphere_run_integrals: null rof buffer for point " + std::to_string(k));
        return FDMGPU_EINVAL;
      }
  }
  bool want_irr = false;
  for (int k = 0; k < count; k++) want_irr = want_irr || P.phi2[k] != nullptr;
  if (want_irr) {
    for (int k = 0; k < count; k++)
      if (!P.phi2[k] && nphi > 0) {
        set_error("fdmgpu_sphere_run_ri: null irregular phiato buffer for point " + std::to_string(k));
        return FDMGPU_EINVAL;
      }
    for (int a = 0; a < natome; a++)
      if (h_inr[a] > h_nr[a] - 1) {
        set_error("fdmgpu_sphere_run_ri: atom " + std::to_string(a) +
                  " needs a radial point beyond its matching point for the irregular solution");
        return FDMGPU_EINVAL;
      }
    if (!u2) FDM_HIP(hipMalloc(&u2, sizeof(double) * (size_t)max_points * natome * (lmax_all + 1) * ustride));
  }
  const int NT = 64;
  const int nth = natome * (lmax_all + 1);
  hipLaunchKernelGGL(k_sph_radial, dim3((nth + NT - 1) / NT, count), dim3(NT), 0, stream, d, P, lmax_all, u, ustride,
                     u2);
  const int nlm_max = (lmax_all + 1) * (lmax_all + 1);
  const int64_t nph = (int64_t)nball * nlm_max;
  if (nph > 0)
    hipLaunchKernelGGL(k_sph_phiato, dim3((unsigned)((nph + 255) / 256), count), dim3(256), 0, stream, d, P, lmax_all,
                       nlm_max, u, ustride, (int)cplx, (const double *)u2);
  if (want_rof)
    hipLaunchKernelGGL(k_sph_rof, dim3((unsigned)nrof(), count), dim3(64), 0, stream, d, P, rt, lmax_all, u, ustride);
  FDM_HIP(hipGetLastError());
  return 0;
}

}  // namespace fdmgpu
