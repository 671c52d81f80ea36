/*
 * ORACLE — test infrastructure only (see fdm_oracle.h).
 *
 * Plain-C restatement of the FDMNES finite-difference path, 1-based like the
 * reference so each loop reads against its Fortran source:
 *   reseau    general.f90:9439-9587   grid points, radial exchange sort, atom points
 *   laplac    general.f90:9732-9871   stencil coefficients, neighbour table
 *   bordure   general.f90:9895-10076  outer-sphere and atom border lists, weights
 *   Ylmpt     general.f90:11505-11570, cylm :11606-11688, ylmcr :11692-11720
 *   phiso     mat.f90:739-1028 (non-Rydberg branch :975-994)
 *   cbessneu  sphere.f90:7226-7248, cbessneur sphere.f90:7356-7378
 *   newind    mat.f90:1761-1921 (atom blocks :1790-1810, :1836-1860)
 *   calcMatRow mat.f90:1034-1632 (grid rows :1142-1398, atom neighbours
 *             :1233-1309, outer rows :1401-1511, atom rows :1515-1632;
 *             scalar-relativistic terms :1155-1156, :1171-1185)
 *   mat_solve mat_solve_gaussian.f90:67-358
 * Constants follow the reference's module declarations (main.f90:36-83).
 */
#define _POSIX_C_SOURCE 199309L
#include "fdm_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* main.f90:55-83 */
static const double PI = 3.1415926535897932384626433832;
static const double SQRT_2 = 1.41421356237309504880168872420970;
static const double EPS6 = 1.e-6;
static const double EPS10 = 1.e-10;
static const double EPS15 = 1.e-15;
static const double EPSPOS = 1.e-4;
static const double ALFA_SF = 0.0072973525664; /* fine-structure constant, main.f90:58 */

struct orc_cluster {
  double rsort, adimp;
  int iord, lmaxso, nvois, nx;
  int npso, npoint, nligne, nligneso, nlmso, nsortf, nsort;
  double *xyz;     /* xyz(4,npso): index 4*(i-1)+k-1 */
  int *indice;     /* indice(npso,3) */
  int *numia;      /* numia(npso) */
  double *rvol;    /* rvol(npoint) */
  int *ivois;      /* ivois(npso,nvois) */
  double clapl[13];
  int *isrt;       /* isrt(nsort) */
  double *poidso;  /* poidso(nsortf) */
  double *ylmso;   /* Ylmso(nsort, nlmso) */
  int *lso, *mso, *iso; /* (nlmso) */
  int *newp;       /* new(0:npoint) */
  int *lb1, *lb2, *newinv; /* (nligne) */
  /* atoms (muffin-tin spheres), point group P1, one spin, regular solution */
  int natome, npr, nbtm, nlmmax;
  double *posi;    /* posi(3,natome), bohr, absorber at the origin */
  double *rmt;     /* (natome) */
  int *lmaxa, *nlmsa, *ianew, *nbord, *nbordf; /* (natome) */
  int *ibord;      /* ibord(nbtm,natome) */
  double *poidsa;  /* poidsa(nbtm,natome), the first nbordf(ia) used */
  double *ylmato;  /* Ylmato(nbtm,nlmmax,natome) */
  double *phi_re, *phi_im; /* phiato(nbtm,nlmmax,natome) set by orc_cluster_set_phiato */
  int phi_cplx;    /* nphiato7 == 2 */
  /* irregular cluster solution (mat.f90:140-147, Test_reg = 0): n_ato_RI = 2
   * blocks per atom, no outer-sphere rows; phi2 = the irregular atomic
   * solutions (i_RI = 2 of phiato) */
  int irregular, n_ato_RI;
  double *phi2_re, *phi2_im;
  /* scalar-relativistic rows (Relativiste, mat.f90:1155-1156, :1171-1185):
   * gradient stencil cgrad(1:nvois) (laplac) and the potential's gradient
   * gradvr(npoint, 3) set by orc_cluster_set_relativistic */
  double cgrad[13];
  int relativiste;
  double *gradvr;
  /* complex harmonics in the outer sphere (Ylm_comp, mat.f90:1329, :1374,
   * :1430-1436, :1476-1497), set by orc_cluster_set_ylm_comp */
  int ylm_comp;
};
#define GRADVR(c, i, k) ((c)->gradvr[((k)-1) * (size_t)(c)->npoint + (i)-1])
#define IBORD(c, ib, ia) ((c)->ibord[((ia)-1) * (c)->nbtm + (ib)-1])
#define POIDSA(c, ib, ia) ((c)->poidsa[((ia)-1) * (c)->nbtm + (ib)-1])
#define YLMATO(c, ib, lm, ia) ((c)->ylmato[(((size_t)(ia)-1) * (c)->nlmmax + (lm)-1) * (c)->nbtm + (ib)-1])
#define PHIATO(a, c, ib, lm, ia) ((a)[(((size_t)(ia)-1) * (c)->nlmmax + (lm)-1) * (c)->nbtm + (ib)-1])

#define XYZ(c, k, i) ((c)->xyz[4 * ((i)-1) + (k)-1])
#define IND(c, i, k) ((c)->indice[((k)-1) * (c)->npso + (i)-1])
#define IVOIS(c, i, iv) ((c)->ivois[((iv)-1) * (c)->npso + (i)-1])
#define YLMSO(c, ib, lm) ((c)->ylmso[((lm)-1) * (c)->nsort + (ib)-1])

static double cpu_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_THREAD_CPUTIME_ID, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* ---------------------------------------------------------------- reseau */
/* general.f90:9439-9561 for a cubic grid, point group P1, Moy_loc = Green = F. */
static int orc_reseau(orc_cluster *c) {
  const double rmax = c->rsort / c->adimp + SQRT_2 * c->iord / 2 + EPSPOS; /* :9476 */
  const int nx = (int)ceil(rmax) + 1;
  c->nx = nx;
  int cap = 0;
  for (int ix = -nx; ix <= nx; ix++)
    for (int iy = -nx; iy <= nx; iy++)
      for (int iz = -nx; iz <= nx; iz++) {
        double v1 = ix, v2 = iy, v3 = iz;
        if (sqrt(v1 * v1 + v2 * v2 + v3 * v3) <= rmax) cap++;
      }
  c->npso = cap;
  c->xyz = calloc((size_t)4 * cap, sizeof(double));
  c->indice = calloc((size_t)3 * cap, sizeof(int));
  c->numia = calloc((size_t)cap + 1, sizeof(int));
  if (!c->xyz || !c->indice || !c->numia) return -1;
  int i = 0;
  for (int ix = -nx; ix <= nx; ix++) /* :9480-9514 loop order ix, iy, iz */
    for (int iy = -nx; iy <= nx; iy++)
      for (int iz = -nx; iz <= nx; iz++) {
        double v1 = ix, v2 = iy, v3 = iz;
        double dist = sqrt(v1 * v1 + v2 * v2 + v3 * v3);
        if (dist > rmax) continue;
        i++;
        IND(c, i, 1) = ix;
        IND(c, i, 2) = iy;
        IND(c, i, 3) = iz;
        XYZ(c, 4, i) = c->adimp * dist;
      }
  const int npso = i;
  /* Exchange sort by distance with tolerance epspos, :9528-9543 */
  for (int a = 1; a <= npso; a++) {
    for (int b = a + 1; b <= npso; b++) {
      if (XYZ(c, 4, a) < XYZ(c, 4, b) + EPSPOS) continue;
      double t = XYZ(c, 4, a);
      XYZ(c, 4, a) = XYZ(c, 4, b);
      XYZ(c, 4, b) = t;
      for (int k = 1; k <= 3; k++) {
        int ti = IND(c, a, k);
        IND(c, a, k) = IND(c, b, k);
        IND(c, b, k) = ti;
      }
    }
  }
  for (int a = 1; a <= npso; a++) /* :9545-9553 */
    for (int k = 1; k <= 3; k++) XYZ(c, k, a) = IND(c, a, k) * c->adimp;
  int a;
  for (a = 1; a <= npso; a++) /* :9556-9559 */
    if (XYZ(c, 4, a) > c->rsort + EPSPOS) break;
  c->npoint = a - 1;
  for (int p = 1; p <= npso; p++) c->numia[p - 1] = (p <= c->npoint) ? 0 : -2;
  /* rvol: with P1 every point is counted once and rvmax = 1, :9582-9651 */
  c->rvol = malloc(sizeof(double) * (c->npoint + 1));
  for (int p = 1; p <= c->npoint; p++) c->rvol[p - 1] = 1.0;
  /* points inside an atom's sphere, :9563-9577 (posrel with the identity
   * only: the distance to the atom) */
  for (int ia = 1; ia <= c->natome; ia++)
    for (int i = 1; i <= c->npoint; i++) {
      double u[3] = {XYZ(c, 1, i) - c->posi[3 * (ia - 1)], XYZ(c, 2, i) - c->posi[3 * (ia - 1) + 1],
                     XYZ(c, 3, i) - c->posi[3 * (ia - 1) + 2]};
      double dr = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
      if (dr < c->rmt[ia - 1] + EPSPOS) c->numia[i - 1] = ia;
    }
  c->npr = c->npoint;
  for (int i = 1; i <= c->npoint; i++)
    if (c->numia[i - 1] != 0) c->npr--;
  return 0;
}

/* ---------------------------------------------------------------- laplac */
/* general.f90:9732-9871, cubic grid. */
static int orc_laplac(orc_cluster *c) {
  const int nx = c->nx, w = 2 * nx + 1;
  int indpro[13][3];
  memset(indpro, 0, sizeof(indpro));
  double ad2 = c->adimp * c->adimp;
  double coeford[2] = {1.0, 0.0};
  double cl;
  if (c->iord == 2) {
    cl = 1.0 / ad2;
  } else {
    cl = (4.0 / 3.0) / ad2;
    coeford[1] = -1.0 / 16;
  }
  int iv = 0;
  for (int io = 1; io <= c->iord / 2; io++)
    for (int i = 1; i <= 3; i++)
      for (int is = -1; is <= 1; is += 2) {
        iv++;
        c->clapl[iv] = coeford[io - 1] * cl;
        indpro[iv][i - 1] = is * io;
      }
  c->nvois = iv;
  double s = 0;
  for (int k = 1; k <= c->nvois; k++) s += c->clapl[k];
  c->clapl[0] = -s;
  /* gradient coefficients, general.f90:9791-9822 (cubic grid) */
  double cd;
  coeford[0] = 1.0;
  if (c->iord == 2) {
    cd = 1 / (2 * c->adimp);
  } else {
    cd = (4.0 / 3.0) / (2 * c->adimp);
    coeford[1] = -1.0 / 8;
  }
  iv = 0;
  for (int io = 1; io <= c->iord / 2; io++)
    for (int i = 1; i <= 3; i++)
      for (int is = -1; is <= 1; is += 2) {
        iv++;
        c->cgrad[iv] = is * coeford[io - 1] * cd;
      }
  /* mpres(-nx:nx,...) */
  int *mpres = calloc((size_t)w * w * w, sizeof(int));
  if (!mpres) return -1;
  for (int p = 1; p <= c->npso; p++)
    mpres[((IND(c, p, 1) + nx) * w + (IND(c, p, 2) + nx)) * w + IND(c, p, 3) + nx] = p;
  c->ivois = calloc((size_t)c->npso * c->nvois, sizeof(int));
  for (int p = 1; p <= c->npso; p++)
    for (int k = 1; k <= c->nvois; k++) {
      int la[3];
      for (int d = 0; d < 3; d++) la[d] = IND(c, p, d + 1) + indpro[k][d];
      if (abs(la[0]) <= nx && abs(la[1]) <= nx && abs(la[2]) <= nx)
        IVOIS(c, p, k) = mpres[((la[0] + nx) * w + (la[1] + nx)) * w + la[2] + nx];
      else
        IVOIS(c, p, k) = 0;
    }
  free(mpres);
  return 0;
}

/* ---------------------------------------------------------------- bordure */
/* general.f90:9895-10076, Green = F, point group P1 (isbord = 1). */
static int orc_bordure(orc_cluster *c) {
  const int nv = 2 * c->nvois / c->iord, na = c->natome;
  int nsortf = 0, iss = 0;
  int *ibp = calloc((size_t)na + 1, sizeof(int));
  c->nbordf = calloc((size_t)na + 1, sizeof(int));
  c->nbord = calloc((size_t)na + 1, sizeof(int));
  if (!ibp || !c->nbordf || !c->nbord) return -1;
  for (int i = 1; i <= c->npoint; i++) { /* :9924-9938 */
    if (c->numia[i - 1] != 0) continue;
    for (int iv = 1; iv <= nv; iv++) {
      int j = IVOIS(c, i, iv);
      if (j == 0) continue;
      int ia = c->numia[j - 1];
      if (ia == -2) {
        if (iss != i) nsortf++;
        iss = i;
      } else if (ia > 0) {
        if (ibp[ia - 1] != i) c->nbordf[ia - 1]++;
        ibp[ia - 1] = i;
      }
    }
  }
  for (int ia = 1; ia <= na; ia++) c->nbord[ia - 1] = c->nbordf[ia - 1];
  for (int i = 1; i <= c->npoint; i++) { /* :9940-9959 */
    int ia = c->numia[i - 1];
    if (ia < 1) continue;
    for (int iv = 1; iv <= c->nvois; iv++) {
      int j = IVOIS(c, i, iv);
      if (j == 0) return -1; /* the reference stops */
      if (c->numia[j - 1] != 0) continue;
      c->nbord[ia - 1]++;
      break;
    }
  }
  int nsort = nsortf;
  for (int i = c->npoint + 1; i <= c->npso; i++) /* :9961-9969 */
    for (int iv = 1; iv <= c->nvois; iv++) {
      int j = IVOIS(c, i, iv);
      if (j == 0) continue;
      if (c->numia[j - 1] != 0) continue;
      nsort++;
      break;
    }
  c->nbtm = 1;
  for (int ia = 1; ia <= na; ia++)
    if (c->nbord[ia - 1] > c->nbtm) c->nbtm = c->nbord[ia - 1];
  c->isrt = calloc((size_t)nsort + 1, sizeof(int));
  c->poidso = calloc((size_t)nsortf + 1, sizeof(double));
  c->ibord = calloc((size_t)c->nbtm * (na > 0 ? na : 1), sizeof(int));
  c->poidsa = calloc((size_t)c->nbtm * (na > 0 ? na : 1), sizeof(double));
  double *pdharm = calloc((size_t)na + 1, sizeof(double));
  if (!c->isrt || !c->poidso || !c->ibord || !c->poidsa || !pdharm) return -1;
  /* second pass, :9987-10039 (iss is not reset, as in the reference) */
  nsortf = 0;
  double pdsort = 0;
  for (int ia = 1; ia <= na; ia++) {
    c->nbordf[ia - 1] = 0;
    ibp[ia - 1] = 0;
  }
  for (int i = 1; i <= c->npoint; i++) {
    if (c->numia[i - 1] != 0) continue;
    for (int iv = 1; iv <= nv; iv++) {
      int j = IVOIS(c, i, iv);
      if (j == 0) continue;
      int ia = c->numia[j - 1];
      if (ia == -2) {
        if (iss != i) {
          pdsort += c->rvol[i - 1];
          nsortf++;
          c->isrt[nsortf - 1] = i;
        }
        iss = i;
      } else if (ia > 0) {
        if (ibp[ia - 1] != i) {
          pdharm[ia - 1] += c->rvol[i - 1];
          c->nbordf[ia - 1]++;
          IBORD(c, c->nbordf[ia - 1], ia) = i;
        }
        ibp[ia - 1] = i;
      }
    }
  }
  for (int ia = 1; ia <= na; ia++) c->nbord[ia - 1] = c->nbordf[ia - 1];
  for (int i = 1; i <= c->npoint; i++) { /* :10015-10039 */
    int ia = c->numia[i - 1];
    if (ia < 1) continue;
    for (int iv = 1; iv <= c->nvois; iv++) {
      int j = IVOIS(c, i, iv);
      if (c->numia[j - 1] != 0) continue;
      c->nbord[ia - 1]++;
      IBORD(c, c->nbord[ia - 1], ia) = i;
      break;
    }
  }
  nsort = nsortf;
  for (int i = c->npoint + 1; i <= c->npso; i++) /* :10041-10050 */
    for (int iv = 1; iv <= c->nvois; iv++) {
      int j = IVOIS(c, i, iv);
      if (j == 0) continue;
      if (c->numia[j - 1] != 0) continue;
      nsort++;
      c->isrt[nsort - 1] = i;
      break;
    }
  for (int ia = 1; ia <= na; ia++) { /* :10063-10068 */
    if (c->nbordf[ia - 1] == 0) continue;
    double fac = 4 * PI / pdharm[ia - 1];
    for (int ib = 1; ib <= c->nbordf[ia - 1]; ib++) POIDSA(c, ib, ia) = fac * c->rvol[IBORD(c, ib, ia) - 1];
  }
  if (nsortf > 0) { /* :10070-10076 */
    double fac = 4 * PI / pdsort;
    for (int ib = 1; ib <= nsortf; ib++) c->poidso[ib - 1] = fac * c->rvol[c->isrt[ib - 1] - 1];
  }
  c->nsortf = nsortf;
  c->nsort = nsort;
  free(ibp);
  free(pdharm);
  return 0;
}

/* ---------------------------------------------------------------- Ylm */
/* cmplx(a, b) / d with d real: Fortran promotes d to complex and the
 * reference build (flang -O2) divides as (a c + b e, b c - a e) / (c^2 + e^2)
 * with c = d, e = 0, two roundings where a/d, b/d would take one. */
static double complex fdiv_by_real(double a, double b, double d) {
  const double den = d * d + 0.0 * 0.0;
  return CMPLX((a * d + b * 0.0) / den, (b * d - a * 0.0) / den);
}

/* cylm, general.f90:11606-11688: complex Y(L,m), m >= 0, index L(L+1)/2+1+m. */
static void orc_cylm(int lmax, const double *v, double r, double complex *ylmc) {
  double cos_t, sin_t = 0, cot_t = 0;
  double complex exphi, sin_t_exp_phi;
  cos_t = (r > EPS6) ? v[2] / r : 1.0;
  double den = sqrt(v[0] * v[0] + v[1] * v[1]);
  if (den > EPS6) {
    exphi = fdiv_by_real(v[0], v[1], den);
    sin_t = sqrt(1 - cos_t * cos_t);
    sin_t_exp_phi = fdiv_by_real(v[0], v[1], r);
    cot_t = cos_t / sin_t;
  } else {
    exphi = 1.0;
    sin_t = 0;
    sin_t_exp_phi = 0;
  }
  ylmc[0] = 1.0 / sqrt(4 * PI);
  for (int L = 1; L <= lmax; L++) { /* Y(L,L) */
    int lm = ((L + 1) * (L + 2)) / 2, lm1 = (L * (L + 1)) / 2;
    double f = -sqrt(1 + 0.5 / L);
    ylmc[lm - 1] = f * sin_t_exp_phi * ylmc[lm1 - 1];
  }
  if (lmax > 0) ylmc[1] = sqrt(3.0 / (4 * PI)) * cos_t; /* Y(1,0) */
  for (int L = 2; L <= lmax; L++) { /* Y(L,L-1) */
    int lm = ((L + 1) * (L + 2)) / 2 - 1, lm1 = (L * L + L) / 2 - 1;
    double f = -sqrt((2 * L + 1.0) / (2 * L - 2));
    ylmc[lm - 1] = f * sin_t_exp_phi * ylmc[lm1 - 1];
  }
  exphi = conj(exphi);
  for (int L = 2; L <= lmax; L++) { /* Y(L,m), L-2 >= m >= 1 */
    int lm0 = (L * L + L) / 2;
    for (int m = L - 2; m >= 1; m--) {
      int lm = lm0 + m + 1;
      double f = -2 * (m + 1.0) / sqrt(L * (L + 1.0) - m * (m + 1.0));
      double g = -sqrt((L * (L + 1.0) - (m + 1.0) * (m + 2.0)) / (L * (L + 1.0) - m * (m + 1.0)));
      if (fabs(sin_t) > EPS10)
        ylmc[lm - 1] = exphi * (cot_t * f * ylmc[lm] + exphi * g * ylmc[lm + 1]);
      else
        ylmc[lm - 1] = 0;
    }
  }
  for (int L = 2; L <= lmax; L++) { /* Y(L,0) */
    int lm = (L * (L + 1)) / 2 + 1, lm1 = ((L - 1) * L) / 2 + 1, lm2 = ((L - 2) * (L - 1)) / 2 + 1;
    double f = sqrt((2 * L + 1.0) / (2 * L - 1.0)) * (2 * L - 1.0) / L;
    double g = -sqrt((2 * L + 1.0) / (2 * L - 3.0)) * (L - 1.0) / L;
    ylmc[lm - 1] = f * cos_t * ylmc[lm1 - 1] + g * ylmc[lm2 - 1];
  }
}

/* ylmcr, general.f90:11692-11720 */
void orc_ylm_real(int lmax, const double *v, double r, double *ylmr) {
  int nlmc = ((lmax + 1) * (lmax + 2)) / 2;
  double complex *ylmc = malloc(sizeof(double complex) * nlmc);
  orc_cylm(lmax, v, r, ylmc);
  int lm = 0;
  for (int L = 0; L <= lmax; L++) {
    int l2 = (L * L + L) / 2 + 1;
    for (int m = -L; m <= L; m++) {
      lm++;
      int lm0 = l2 + abs(m);
      double sgn = (abs(m) % 2) ? -1.0 : 1.0;
      if (m < 0)
        ylmr[lm - 1] = sgn * SQRT_2 * cimag(ylmc[lm0 - 1]);
      else if (m == 0)
        ylmr[lm - 1] = creal(ylmc[lm0 - 1]);
      else
        ylmr[lm - 1] = sgn * SQRT_2 * creal(ylmc[lm0 - 1]);
    }
  }
  free(ylmc);
}

/* Ylmpt for the outer sphere (ia = 0), general.f90:11545-11568 */
static void orc_ylmpt(orc_cluster *c) {
  const int nlmr = (c->lmaxso + 1) * (c->lmaxso + 1);
  c->ylmso = calloc((size_t)c->nsort * nlmr, sizeof(double));
  double *y = malloc(sizeof(double) * nlmr);
  for (int ib = 1; ib <= c->nsort; ib++) {
    int i = c->isrt[ib - 1];
    double w[3] = {XYZ(c, 1, i), XYZ(c, 2, i), XYZ(c, 3, i)};
    double r = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    orc_ylm_real(c->lmaxso, w, r, y);
    for (int lm = 1; lm <= nlmr; lm++) YLMSO(c, ib, lm) = y[lm - 1];
  }
  free(y);
  /* the atoms (ia > 0), :11533-11568: every border point, its direction
   * from the atom (posrel with the identity, rot_atom = 1) */
  c->nlmmax = 1;
  for (int ia = 1; ia <= c->natome; ia++)
    if (c->nlmsa[ia - 1] > c->nlmmax) c->nlmmax = c->nlmsa[ia - 1];
  c->ylmato = calloc((size_t)c->nbtm * c->nlmmax * (c->natome > 0 ? c->natome : 1), sizeof(double));
  y = malloc(sizeof(double) * c->nlmmax);
  for (int ia = 1; ia <= c->natome; ia++)
    for (int ib = 1; ib <= c->nbord[ia - 1]; ib++) {
      int i = IBORD(c, ib, ia);
      double v[3] = {XYZ(c, 1, i) - c->posi[3 * (ia - 1)], XYZ(c, 2, i) - c->posi[3 * (ia - 1) + 1],
                     XYZ(c, 3, i) - c->posi[3 * (ia - 1) + 2]};
      double r = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      orc_ylm_real(c->lmaxa[ia - 1], v, r, y);
      for (int lm = 1; lm <= c->nlmsa[ia - 1]; lm++) YLMATO(c, ib, lm, ia) = y[lm - 1];
    }
  free(y);
}

/* ---------------------------------------------------------------- newind */
/* mat.f90:1761-1921, nspino = 1, regular solution (n_ato_RI = 1); the atom
 * blocks are placed in the radial order at their distance distai(ia) from
 * the absorber (:1790-1810). */
static void orc_newind(orc_cluster *c) {
  const int nspino = 1, na = c->natome, nri = c->n_ato_RI;
  c->newp = calloc((size_t)c->npoint + 2, sizeof(int));
  c->ianew = calloc((size_t)na + 1, sizeof(int));
  int ii = 0;
  for (int i = 1; i <= c->npoint; i++) {
    if (c->numia[i - 1] == 0) {
      ii += nspino;
      c->newp[i] = ii;
    } else {
      c->newp[i] = 0;
    }
  }
  int npf = 1, npf1 = 0;
  for (int ia = 1; ia <= na; ia++) { /* :1790-1810 */
    const double *p = c->posi + 3 * (ia - 1);
    const double distai = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    int i;
    for (i = npf; i <= c->npoint; i++)
      if (c->numia[i - 1] == 0 && XYZ(c, 4, i) >= distai) break;
    npf = i;
    if (npf1 == npf)
      c->ianew[ia - 1] = c->ianew[ia - 2] + nri * c->nlmsa[ia - 2];
    else
      c->ianew[ia - 1] = c->newp[npf] - nspino;
    npf1 = npf;
    for (i = npf; i <= c->npoint; i++)
      if (c->numia[i - 1] == 0) c->newp[i] += nri * c->nlmsa[ia - 1];
  }
  const int nligne = c->nligne;
  c->lb1 = malloc(sizeof(int) * (nligne + 1));
  c->lb2 = malloc(sizeof(int) * (nligne + 1));
  c->newinv = calloc((size_t)nligne + 1, sizeof(int));
  for (ii = 1; ii <= nligne; ii++) {
    c->lb1[ii] = ii;
    c->lb2[ii] = ii;
  }
  const int n = nspino - 1;
  for (int i = 1; i <= c->npoint; i++) { /* :1820-1861 */
    if (c->numia[i - 1] != 0) continue;
    ii = c->newp[i];
    for (int iv = 1; iv <= c->nvois; iv++) {
      int j = IVOIS(c, i, iv);
      int ia = c->numia[j - 1];
      if (ia == 0) {
        int nj = c->newp[j];
        if (nj - n < c->lb1[ii]) c->lb1[ii] = nj - n;
        if (nj > c->lb2[ii]) c->lb2[ii] = nj;
      } else if (ia > 0) {
        if (c->ianew[ia - 1] + 1 < c->lb1[ii]) c->lb1[ii] = c->ianew[ia - 1] + 1;
        if (c->ianew[ia - 1] + nri * c->nlmsa[ia - 1] > c->lb2[ii]) c->lb2[ii] = c->ianew[ia - 1] + nri * c->nlmsa[ia - 1];
      } else if (ia == -2) {
        c->lb2[ii] = nligne;
      }
    }
  }
  for (int ia = 1; ia <= na; ia++) { /* :1863-1875 */
    int lb11 = c->ianew[ia - 1] + 1, lb22 = c->ianew[ia - 1] + nri * c->nlmsa[ia - 1];
    for (int ib = 1; ib <= c->nbordf[ia - 1]; ib++) {
      int r = c->newp[IBORD(c, ib, ia)];
      if (r - n < lb11) lb11 = r - n;
      if (r > lb22) lb22 = r;
    }
    for (int L = 1; L <= nri * c->nlmsa[ia - 1]; L++) {
      ii = c->ianew[ia - 1] + L;
      c->lb1[ii] = lb11;
      c->lb2[ii] = lb22;
    }
  }
  if (!c->irregular) { /* :1877-1886, no scattered amplitude in the irregular solution */
    int lb11 = nligne - c->nlmso + 1;
    for (int ib = 1; ib <= c->nsortf; ib++) {
      int r = c->newp[c->isrt[ib - 1]] - n;
      if (r < lb11) lb11 = r;
    }
    for (ii = nligne - c->nlmso + 1; ii <= nligne; ii++) {
      c->lb1[ii] = lb11;
      c->lb2[ii] = nligne;
    }
  }
  for (ii = nligne - 1; ii >= 1; ii--) /* :1888-1890 */
    if (c->lb1[ii + 1] < c->lb1[ii]) c->lb1[ii] = c->lb1[ii + 1];
  for (int i = 1; i <= c->npoint; i++) {
    ii = c->newp[i];
    if (ii != 0) c->newinv[ii] = i;
  }
  for (int ia = 1; ia <= na; ia++)
    for (int Lm = 1; Lm <= nri * c->nlmsa[ia - 1]; Lm++) c->newinv[c->ianew[ia - 1] + Lm] = -ia;
}

orc_cluster *orc_cluster_new_atoms_ri(double rsort, double adimp, int iord, int lmaxso, int natome,
                                      const double *posi, const double *rmt, const int *lmaxa, int irregular) {
  orc_cluster *c = calloc(1, sizeof(orc_cluster));
  if (!c) return NULL;
  c->irregular = irregular != 0;
  c->n_ato_RI = c->irregular ? 2 : 1;
  c->rsort = rsort;
  c->adimp = adimp;
  c->iord = iord;
  c->lmaxso = lmaxso;
  c->natome = natome;
  c->posi = calloc((size_t)3 * natome + 1, sizeof(double));
  c->rmt = calloc((size_t)natome + 1, sizeof(double));
  c->lmaxa = calloc((size_t)natome + 1, sizeof(int));
  c->nlmsa = calloc((size_t)natome + 1, sizeof(int));
  if (!c->posi || !c->rmt || !c->lmaxa || !c->nlmsa) {
    orc_cluster_free(c);
    return NULL;
  }
  for (int ia = 0; ia < natome; ia++) {
    for (int k = 0; k < 3; k++) c->posi[3 * ia + k] = posi[3 * ia + k];
    c->rmt[ia] = rmt[ia];
    c->lmaxa[ia] = lmaxa[ia];
    c->nlmsa[ia] = (lmaxa[ia] + 1) * (lmaxa[ia] + 1); /* one spin: all (L, m) of L <= lmax */
  }
  if (orc_reseau(c) || orc_laplac(c) || orc_bordure(c)) {
    orc_cluster_free(c);
    return NULL;
  }
  orc_ylmpt(c);
  /* outer harmonics with P1: Lm = 1..(lmaxso+1)^2 in (L, m) order, iso = 1 */
  c->nlmso = (lmaxso + 1) * (lmaxso + 1);
  c->lso = malloc(sizeof(int) * (c->nlmso + 1));
  c->mso = malloc(sizeof(int) * (c->nlmso + 1));
  c->iso = malloc(sizeof(int) * (c->nlmso + 1));
  int lm = 0;
  for (int L = 0; L <= lmaxso; L++)
    for (int m = -L; m <= L; m++) {
      lm++;
      c->lso[lm] = L;
      c->mso[lm] = m;
      c->iso[lm] = 1;
    }
  int nat = 0; /* mat.f90:147, :150 */
  for (int ia = 0; ia < natome; ia++) nat += c->nlmsa[ia];
  if (c->irregular) {
    c->nligne = c->npr + c->n_ato_RI * nat;
    c->nligneso = c->nligne; /* mat.f90:178-179 */
  } else {
    c->nligne = c->npr + nat + c->nlmso;
    c->nligneso = c->nligne - c->nlmso;
  }
  orc_newind(c);
  return c;
}

orc_cluster *orc_cluster_new_atoms(double rsort, double adimp, int iord, int lmaxso, int natome, const double *posi,
                                   const double *rmt, const int *lmaxa) {
  return orc_cluster_new_atoms_ri(rsort, adimp, iord, lmaxso, natome, posi, rmt, lmaxa, 0);
}

orc_cluster *orc_cluster_new(double rsort, double adimp, int iord, int lmaxso) {
  return orc_cluster_new_atoms(rsort, adimp, iord, lmaxso, 0, NULL, NULL, NULL);
}

void orc_cluster_free(orc_cluster *c) {
  if (!c) return;
  free(c->xyz); free(c->indice); free(c->numia); free(c->rvol); free(c->ivois);
  free(c->isrt); free(c->poidso); free(c->ylmso); free(c->lso); free(c->mso); free(c->iso);
  free(c->newp); free(c->lb1); free(c->lb2); free(c->newinv);
  free(c->gradvr);
  free(c->posi); free(c->rmt); free(c->lmaxa); free(c->nlmsa); free(c->ianew); free(c->nbord); free(c->nbordf);
  free(c->ibord); free(c->poidsa); free(c->ylmato); free(c->phi_re); free(c->phi_im);
  free(c->phi2_re); free(c->phi2_im);
  free(c);
}

void orc_cluster_sizes(const orc_cluster *c, int *s) {
  s[0] = c->npoint; s[1] = c->npso; s[2] = c->nligne; s[3] = c->nlmso;
  s[4] = c->nsortf; s[5] = c->nsort; s[6] = c->nvois;
}
void orc_cluster_atom_sizes(const orc_cluster *c, int *s) {
  s[0] = c->natome; s[1] = c->npr; s[2] = c->nbtm; s[3] = c->nlmmax;
}
void orc_cluster_atom(const orc_cluster *c, int ia, int *info, int *ibord, double *poidsa, double *ylmato) {
  info[0] = c->nbord[ia - 1]; info[1] = c->nbordf[ia - 1]; info[2] = c->nlmsa[ia - 1]; info[3] = c->ianew[ia - 1];
  for (int ib = 1; ib <= c->nbord[ia - 1]; ib++) {
    if (ibord) ibord[ib - 1] = IBORD(c, ib, ia);
    if (poidsa) poidsa[ib - 1] = ib <= c->nbordf[ia - 1] ? POIDSA(c, ib, ia) : 0.0;
    if (ylmato)
      for (int lm = 1; lm <= c->nlmsa[ia - 1]; lm++)
        ylmato[(size_t)(ib - 1) * c->nlmsa[ia - 1] + lm - 1] = YLMATO(c, ib, lm, ia);
  }
}
void orc_cluster_numia(const orc_cluster *c, int *numia) { memcpy(numia, c->numia, sizeof(int) * c->npso); }
/* i_RI = 1: the regular atomic solutions (phi_re/phi_im); 2: the irregular
 * ones (phi2_re/phi2_im) */
int orc_cluster_set_phiato_ri(orc_cluster *c, int i_ri, int cplx, const double *re, const double *im) {
  const size_t n = (size_t)c->nbtm * c->nlmmax * (c->natome > 0 ? c->natome : 1);
  double **pr = i_ri == 2 ? &c->phi2_re : &c->phi_re, **pi = i_ri == 2 ? &c->phi2_im : &c->phi_im;
  free(*pr);
  free(*pi);
  *pr = calloc(n, sizeof(double));
  *pi = calloc(n, sizeof(double));
  if (!*pr || !*pi) return -1;
  if (i_ri != 2) c->phi_cplx = cplx;
  size_t q = 0; /* input: per atom, per border point ib, nlmsa(ia) harmonics */
  for (int ia = 1; ia <= c->natome; ia++)
    for (int ib = 1; ib <= c->nbord[ia - 1]; ib++)
      for (int lm = 1; lm <= c->nlmsa[ia - 1]; lm++, q++) {
        PHIATO(*pr, c, ib, lm, ia) = re[q];
        PHIATO(*pi, c, ib, lm, ia) = (cplx && im) ? im[q] : 0.0;
      }
  return 0;
}
int orc_cluster_set_relativistic(orc_cluster *c, int on, const double *gradvr) {
  free(c->gradvr);
  c->gradvr = NULL;
  c->relativiste = 0;
  if (!on) return 0;
  if (!gradvr) return -1;
  c->gradvr = malloc(sizeof(double) * 3 * (size_t)c->npoint);
  if (!c->gradvr) return -1;
  memcpy(c->gradvr, gradvr, sizeof(double) * 3 * (size_t)c->npoint);
  c->relativiste = 1;
  return 0;
}
int orc_cluster_set_ylm_comp(orc_cluster *c, int on) {
  if (c->natome > 0 && on) return -1; /* the atom branches are restated for real harmonics */
  c->ylm_comp = on != 0;
  return 0;
}
void orc_cluster_cgrad(const orc_cluster *c, double *cg) { memcpy(cg, c->cgrad + 1, sizeof(double) * c->nvois); }
int orc_cluster_set_phiato(orc_cluster *c, int cplx, const double *re, const double *im) {
  return orc_cluster_set_phiato_ri(c, 1, cplx, re, im);
}
void orc_cluster_xyz(const orc_cluster *c, double *xyz) { memcpy(xyz, c->xyz, sizeof(double) * 4 * c->npso); }
void orc_cluster_indice(const orc_cluster *c, int *ind) { memcpy(ind, c->indice, sizeof(int) * 3 * c->npso); }
void orc_cluster_ivois(const orc_cluster *c, int *iv) { memcpy(iv, c->ivois, sizeof(int) * c->npso * c->nvois); }
void orc_cluster_boundary(const orc_cluster *c, int *isrt, double *poidso) {
  memcpy(isrt, c->isrt, sizeof(int) * c->nsort);
  memcpy(poidso, c->poidso, sizeof(double) * c->nsortf);
}
void orc_cluster_ylmso(const orc_cluster *c, double *y) { memcpy(y, c->ylmso, sizeof(double) * c->nsort * c->nlmso); }
void orc_cluster_clapl(const orc_cluster *c, double *cl) { memcpy(cl, c->clapl, sizeof(double) * (c->nvois + 1)); }
void orc_cluster_profile(const orc_cluster *c, int *lb1, int *lb2, int *newinv, int *newp) {
  memcpy(lb1, c->lb1 + 1, sizeof(int) * c->nligne);
  memcpy(lb2, c->lb2 + 1, sizeof(int) * c->nligne);
  if (newinv) memcpy(newinv, c->newinv + 1, sizeof(int) * c->nligne);
  if (newp) memcpy(newp, c->newp, sizeof(int) * (c->npoint + 1));
}

/* ---------------------------------------------------------------- Bessel */
void orc_cbessneu(double fn_re, double fn_im, double z_re, double z_im, int lmax, double *bes, double *neu) {
  /* sphere.f90:7226-7248; bes/neu hold (re,im) pairs for L = 0..lmax */
  double complex fnorm = CMPLX(fn_re, fn_im), z = CMPLX(z_re, z_im);
  double complex *b = malloc(sizeof(double complex) * (lmax + 1));
  double complex *n = malloc(sizeof(double complex) * (lmax + 1));
  n[0] = -fnorm * ccos(z) / z;
  b[0] = fnorm * csin(z) / z;
  if (lmax >= 1) {
    n[1] = n[0] / z - b[0];
    b[1] = b[0] / z + n[0];
    for (int L = 2; L <= lmax; L++) {
      int L1 = 2 * L - 1;
      n[L] = L1 * n[L - 1] / z - n[L - 2];
      b[L] = L1 * b[L - 1] / z - b[L - 2];
    }
  }
  for (int L = 0; L <= lmax; L++) {
    bes[2 * L] = creal(b[L]); bes[2 * L + 1] = cimag(b[L]);
    neu[2 * L] = creal(n[L]); neu[2 * L + 1] = cimag(n[L]);
  }
  free(b); free(n);
}

void orc_cbessneur(double fnorm, double z, int lmax, double *bes, double *neu) {
  /* sphere.f90:7356-7378 */
  neu[0] = -fnorm * cos(z) / z;
  bes[0] = fnorm * sin(z) / z;
  if (lmax == 0) return;
  neu[1] = neu[0] / z - bes[0];
  bes[1] = bes[0] / z + neu[0];
  for (int L = 2; L <= lmax; L++) {
    int L1 = 2 * L - 1;
    neu[L] = L1 * neu[L - 1] / z - neu[L - 2];
    bes[L] = L1 * bes[L - 1] / z - bes[L - 2];
  }
}

/* phiso non-Rydberg branch, mat.f90:771-792 and :975-994.
 * Complex: bes/neu are (nsort, 0:lmaxso) of (re,im) pairs, ib fastest. */
void orc_phiso(const orc_cluster *c, int E_comp, double Ecinetic_out, double Eimag, double *bes, double *neu) {
  const int lmax = c->lmaxso, ns = c->nsort;
  double *tb = malloc(sizeof(double) * 2 * (lmax + 1)), *tn = malloc(sizeof(double) * 2 * (lmax + 1));
  if (E_comp) {
    double complex konde = csqrt(CMPLX(Ecinetic_out, Eimag));
    if (cabs(konde) < EPS10) konde = CMPLX(EPS10, 0.0);
    double complex fnormc = csqrt(konde / PI);
    for (int ib = 1; ib <= ns; ib++) {
      int i = c->isrt[ib - 1];
      double p1 = XYZ(c, 1, i), p2 = XYZ(c, 2, i), p3 = XYZ(c, 3, i);
      double complex z = konde * sqrt(p1 * p1 + p2 * p2 + p3 * p3);
      orc_cbessneu(creal(fnormc), cimag(fnormc), creal(z), cimag(z), lmax, tb, tn);
      for (int L = 0; L <= lmax; L++) {
        size_t o = 2 * ((size_t)L * ns + ib - 1);
        bes[o] = tb[2 * L]; bes[o + 1] = tb[2 * L + 1];
        neu[o] = tn[2 * L]; neu[o + 1] = tn[2 * L + 1];
      }
    }
  } else {
    double konder = sqrt(Ecinetic_out);
    double fnorm = sqrt(konder / PI);
    for (int ib = 1; ib <= ns; ib++) {
      int i = c->isrt[ib - 1];
      double p1 = XYZ(c, 1, i), p2 = XYZ(c, 2, i), p3 = XYZ(c, 3, i);
      double zr = konder * sqrt(p1 * p1 + p2 * p2 + p3 * p3);
      orc_cbessneur(fnorm, zr, lmax, tb, tn);
      for (int L = 0; L <= lmax; L++) {
        size_t o = (size_t)L * ns + ib - 1;
        bes[o] = tb[L];
        neu[o] = tn[L];
      }
    }
  }
  free(tb); free(tn);
}

/* ---------------------------------------------------------------- calcMatRow */
/* Yc, mat.f90:1733-1755: the complex harmonic (L, m) from the real ones
 * y1 = Y(L, m), y2 = Y(L, -m) */
static const double SQRT_1O2 = 1 / 1.41421356237309504880168872420970;
static double complex orc_yc(int m, double y1, double y2) {
  if (m == 0) return CMPLX(y1, 0.0);
  if (m < 0) return CMPLX(y2 * SQRT_1O2, -y1 * SQRT_1O2);
  const double s = (m & 1) ? -1.0 : 1.0;
  return CMPLX(y1 * s * SQRT_1O2, y2 * s * SQRT_1O2);
}
/* the outer-sphere harmonic of unknown (L, m) at boundary point ib: Yc with
 * complex harmonics, else the real one */
static double complex orc_ysort(const orc_cluster *c, int ib, int lm0, int m) {
  if (c->ylm_comp && m != 0) return orc_yc(m, YLMSO(c, ib, lm0), YLMSO(c, ib, lm0 - 2 * m));
  return CMPLX(YLMSO(c, ib, lm0), 0.0);
}

typedef struct {
  const orc_cluster *c;
  const double *V;
  double Enervide, Eimag;
  int E_comp, cal_comp, basereel;
  const double *bes, *neu; /* complex (E_comp) or real, (nsort, 0:lmaxso) */
} row_ctx;

static inline double complex ctx_bes(const row_ctx *x, int ib, int L) {
  size_t o = (size_t)L * x->c->nsort + ib - 1;
  return x->E_comp ? CMPLX(x->bes[2 * o], x->bes[2 * o + 1]) : CMPLX(x->bes[o], 0.0);
}
static inline double complex ctx_neu(const row_ctx *x, int ib, int L) {
  size_t o = (size_t)L * x->c->nsort + ib - 1;
  return x->E_comp ? CMPLX(x->neu[2 * o], x->neu[2 * o + 1]) : CMPLX(x->neu[o], 0.0);
}

/* mat.f90:1034-1511 restricted to the oracle's branch set.  abvr/abvi are
 * indexed by absolute column (1-based); smr/smi rows are (nlmso) at row ii.
 * abvi may be NULL when cal_comp == 0. */
static void orc_calcmatrow(const row_ctx *x, int ii, double *abvr, double *abvi, double *smr_ii, double *smi_ii) {
  const orc_cluster *c = x->c;
  const int nlmso = c->nlmso, nligneso = c->nligneso;
  const int i = c->newinv[ii];
  const int cal = x->cal_comp;
  if (i > 0) {
    /* Diagonal term, :1147-1150 */
    double vme = x->V[i - 1] - x->Enervide;
    abvr[ii] = vme - c->clapl[0];
    if (x->E_comp) abvi[ii] = -x->Eimag;
    /* :1155-1156 (a2s4 :1123) */
    const double a2s4 = 0.25 * (ALFA_SF * ALFA_SF);
    double bder = 0;
    if (c->relativiste) {
      abvr[ii] = abvr[ii] - a2s4 * (vme * vme);
      bder = -a2s4 / (1 - a2s4 * vme);
    }
    for (int iv = 1; iv <= c->nvois; iv++) {
      int j = IVOIS(c, i, iv);
      int ia = c->numia[j - 1];
      /* :1164-1232 with Kar = 1, Kari = 0, Repres_comp = F */
      double cfvs_r = -c->clapl[iv];
      double cfvs_i = 0.0;
      if (c->relativiste) { /* :1171-1185, cubic grid: k = x, y, z of the pair */
        const int k = 1 + ((iv - 1) / 2) % 3;
        double crelat = bder * c->cgrad[iv] * GRADVR(c, i, k);
        cfvs_r = cfvs_r + crelat;
      }
      cfvs_r = 1.0 * cfvs_r;
      cfvs_r = cfvs_r * c->rvol[i - 1];
      if (cal) cfvs_i = cfvs_i * c->rvol[i - 1];
      if (ia == 0) { /* :1227-1234 */
        int jj = c->newp[j];
        abvr[jj] = abvr[jj] + cfvs_r / c->rvol[j - 1];
        if (cal) abvi[jj] = abvi[jj] + cfvs_i / c->rvol[j - 1];
      } else if (ia > 0) { /* neighbour inside atom ia, :1233-1309 (isym = 1) */
        int ib;
        for (ib = c->nbordf[ia - 1] + 1; ib <= c->nbord[ia - 1]; ib++)
          if (j == IBORD(c, ib, ia)) break;
        /* :1266-1307: regular (i_RI = 1) and, in the irregular solution,
         * irregular (i_RI = 2) atomic solutions, Test_reg = 0 */
        for (int i_ri = 1; i_ri <= c->n_ato_RI; i_ri++) {
          const double *phr = i_ri == 2 ? c->phi2_re : c->phi_re, *phi = i_ri == 2 ? c->phi2_im : c->phi_im;
          for (int Lm = 1; Lm <= c->nlmsa[ia - 1]; Lm++) {
            int jjj = c->ianew[ia - 1] + Lm;
            if (i_ri == 2 && c->n_ato_RI == 2) jjj += c->nlmsa[ia - 1];
            /* Lato/mato in (L, m) order: Lm0 = L**2 + L + 1 + m = Lm */
            const double pr = PHIATO(phr, c, ib, Lm, ia), pi = PHIATO(phi, c, ib, Lm, ia);
            if (c->phi_cplx) {
              double complex cfac = CMPLX(cfvs_r, cfvs_i) * CMPLX(pr, pi);
              abvr[jjj] = abvr[jjj] + creal(cfac);
              abvi[jjj] = abvi[jjj] + cimag(cfac);
            } else if (cal) {
              abvr[jjj] = abvr[jjj] + cfvs_r * pr;
              abvi[jjj] = abvi[jjj] + cfvs_i * pr;
            } else {
              abvr[jjj] = abvr[jjj] + cfvs_r * pr;
            }
          }
        }
      } else if (ia == -2) { /* :1312-1398 */
        int ib;
        for (ib = c->nsortf + 1; ib <= c->nsort; ib++)
          if (j == c->isrt[ib - 1]) break;
        double complex cfvs = CMPLX(cfvs_r, cfvs_i);
        /* :1318-1361: no scattered wave in the irregular solution */
        for (int lm = 1; lm <= (c->irregular ? 0 : nlmso); lm++) {
          int L = c->lso[lm], m = c->mso[lm];
          int lm0 = L * L + L + 1 + m;
          int jjj = nligneso + lm;
          if (cal) {
            double complex ycomp = orc_ysort(c, ib, lm0, m); /* :1328-1332 */
            double complex cfac;
            if (x->basereel) {
              cfac = ctx_neu(x, ib, L);
            } else {
              cfac = ctx_bes(x, ib, L) + I * ctx_neu(x, ib, L);
              if (!x->E_comp) cfac = CMPLX(creal(ctx_bes(x, ib, L)), creal(ctx_neu(x, ib, L)));
            }
            double complex cfv = cfac * ycomp * cfvs;
            abvr[jjj] = abvr[jjj] + creal(cfv);
            abvi[jjj] = abvi[jjj] + cimag(cfv);
          } else {
            double fr = YLMSO(c, ib, lm0) * creal(ctx_neu(x, ib, L));
            abvr[jjj] = abvr[jjj] + cfvs_r * fr;
          }
        }
        for (int lm = 1; lm <= nlmso; lm++) {
          int L = c->lso[lm], m = c->mso[lm];
          int lm0 = L * L + L + 1 + m;
          if (cal) {
            double complex ycomp = orc_ysort(c, ib, lm0, m); /* :1373-1377 */
            double complex cfv;
            if (c->irregular) /* :1379-1381 */
              cfv = ycomp * (-I * ctx_bes(x, ib, L) + ctx_neu(x, ib, L)) * cfvs;
            else
              cfv = ycomp * ctx_bes(x, ib, L) * cfvs;
            smr_ii[lm - 1] = smr_ii[lm - 1] - creal(cfv);
            smi_ii[lm - 1] = smi_ii[lm - 1] - cimag(cfv);
          } else {
            double fr = YLMSO(c, ib, lm0) * creal(ctx_bes(x, ib, L));
            smr_ii[lm - 1] = smr_ii[lm - 1] - cfvs_r * fr;
          }
        }
      }
    }
  } else if (i == 0) {
    /* Outer-sphere rows, :1401-1511 */
    const int lm = ii - nligneso;
    const int m = c->mso[lm];
    const int lm0 = c->lso[lm] * c->lso[lm] + c->lso[lm] + 1 + m;
    const int isp = c->iso[lm];
    for (int lmp = 1; lmp <= nlmso; lmp++) {
      if (c->iso[lmp] != isp) continue;
      int L = c->lso[lmp], mp = c->mso[lmp];
      int lmp0 = L * L + L + 1 + mp;
      int ljj = nligneso + lmp;
      double complex charmc = 0;
      double charmr = 0;
      for (int ib = 1; ib <= c->nsortf; ib++) {
        if (cal) {
          double complex cfac; /* :1429-1437 */
          if (c->ylm_comp && m != 0 && mp != 0)
            cfac = conj(orc_yc(m, YLMSO(c, ib, lm0), YLMSO(c, ib, lm0 - 2 * m))) *
                   orc_yc(mp, YLMSO(c, ib, lmp0), YLMSO(c, ib, lmp0 - 2 * mp));
          else if (c->ylm_comp && m != 0)
            cfac = conj(orc_yc(m, YLMSO(c, ib, lm0), YLMSO(c, ib, lm0 - 2 * m))) * YLMSO(c, ib, lmp0);
          else if (c->ylm_comp && mp != 0)
            cfac = YLMSO(c, ib, lm0) * orc_yc(mp, YLMSO(c, ib, lmp0), YLMSO(c, ib, lmp0 - 2 * mp));
          else
            cfac = CMPLX(YLMSO(c, ib, lm0) * YLMSO(c, ib, lmp0), 0.0);
          cfac = c->poidso[ib - 1] * cfac;
          if (x->basereel) {
            charmc = charmc + cfac * ctx_neu(x, ib, L);
          } else if (x->E_comp) {
            charmc = charmc + cfac * (ctx_bes(x, ib, L) + I * ctx_neu(x, ib, L));
          } else {
            charmc = charmc + cfac * CMPLX(creal(ctx_bes(x, ib, L)), creal(ctx_neu(x, ib, L)));
          }
        } else {
          double fac = c->poidso[ib - 1] * YLMSO(c, ib, lm0) * YLMSO(c, ib, lmp0);
          charmr = charmr + fac * creal(ctx_neu(x, ib, L));
        }
      }
      if (cal) {
        abvr[ljj] = abvr[ljj] + creal(charmc);
        abvi[ljj] = abvi[ljj] + cimag(charmc);
      } else {
        abvr[ljj] = abvr[ljj] + charmr;
      }
    }
    for (int ib = 1; ib <= c->nsortf; ib++) {
      int j = c->isrt[ib - 1];
      int jj = c->newp[j] - 1 + isp;
      double fac = c->poidso[ib - 1] * YLMSO(c, ib, lm0);
      double complex cfac = CMPLX(fac, 0.0);
      if (c->ylm_comp && m != 0) { /* :1474-1478 */
        cfac = c->poidso[ib - 1] * conj(orc_yc(m, YLMSO(c, ib, lm0), YLMSO(c, ib, lm0 - 2 * m)));
        abvr[jj] = abvr[jj] - creal(cfac) / c->rvol[j - 1];
        abvi[jj] = abvi[jj] - cimag(cfac) / c->rvol[j - 1];
      } else
      abvr[jj] = abvr[jj] - fac / c->rvol[j - 1];
      for (int lmf = 1; lmf <= nlmso; lmf++) {
        if (isp != c->iso[lmf]) continue;
        int L = c->lso[lmf], mf = c->mso[lmf];
        int lmf0 = L * L + L + 1 + mf;
        if (cal) {
          double complex cfad = cfac * orc_ysort(c, ib, lmf0, mf); /* :1492-1497 */
          cfad = cfad * (x->E_comp ? ctx_bes(x, ib, L) : CMPLX(creal(ctx_bes(x, ib, L)), 0.0));
          smr_ii[lmf - 1] = smr_ii[lmf - 1] - creal(cfad);
          smi_ii[lmf - 1] = smi_ii[lmf - 1] - cimag(cfad);
        } else {
          double fad = fac * YLMSO(c, ib, lmf0);
          smr_ii[lmf - 1] = smr_ii[lmf - 1] - fad * creal(ctx_bes(x, ib, L));
        }
      }
    }
  } else {
    /* Expansion in atom ia, :1515-1632 (regular solution, Ylm_comp = F,
     * Kar = 1, one spin: ispt = isol = 1) */
    const int ia = -i;
    /* :1528-1545: in the irregular solution rows ianew + 1 .. + nlmsa match
     * the grid on the first half of the border points (i_RI = 1), the next
     * nlmsa rows on the second half (i_RI = 2) */
    int Lm = ii - c->ianew[ia - 1]; /* Lm0 = Lm with (L, m) ordering */
    int i_ri = 1;
    if (c->n_ato_RI == 2 && Lm > c->nlmsa[ia - 1]) {
      i_ri = 2;
      Lm -= c->nlmsa[ia - 1];
    }
    const int nh = c->nbordf[ia - 1] / 2;
    for (int ib = 1; ib <= c->nbordf[ia - 1]; ib++) {
      if (c->irregular && c->n_ato_RI == 2) {
        if (i_ri == 1 && ib > nh) break;
        if (i_ri == 2 && ib <= nh) continue;
      }
      int j = IBORD(c, ib, ia);
      int jj = c->newp[j];
      if (cal) {
        double complex cfac = POIDSA(c, ib, ia) * conj(CMPLX(YLMATO(c, ib, Lm, ia), 0.0)) * 1.0 / c->rvol[j - 1];
        abvr[jj] = abvr[jj] - creal(cfac);
        abvi[jj] = abvi[jj] - cimag(cfac);
      } else {
        double f = POIDSA(c, ib, ia) * 1.0 / c->rvol[j - 1];
        abvr[jj] = abvr[jj] - f * YLMATO(c, ib, Lm, ia);
      }
    }
    for (int j_ri = 1; j_ri <= (c->irregular ? 2 : 1); j_ri++) /* :1580-1630 */
    for (int Lmp = 1; Lmp <= c->nlmsa[ia - 1]; Lmp++) {
      const double *phr = j_ri == 2 ? c->phi2_re : c->phi_re, *phi = j_ri == 2 ? c->phi2_im : c->phi_im;
      double charmr = 0, charmi = 0;
      for (int ib = 1; ib <= c->nbordf[ia - 1]; ib++) {
        if (c->irregular) {
          if (j_ri == 1 && ib > nh) break;
          if (j_ri == 2 && ib <= nh) continue;
        }
        const double pr = PHIATO(phr, c, ib, Lmp, ia), pi = PHIATO(phi, c, ib, Lmp, ia);
        if (cal) {
          double complex cfac = c->phi_cplx ? CMPLX(pr, pi) : CMPLX(pr, 0.0);
          cfac = cfac * YLMATO(c, ib, Lm, ia);
          charmr = charmr + POIDSA(c, ib, ia) * creal(cfac);
          charmi = charmi + POIDSA(c, ib, ia) * cimag(cfac);
        } else {
          charmr = charmr + POIDSA(c, ib, ia) * YLMATO(c, ib, Lm, ia) * pr;
        }
      }
      int ljj = c->ianew[ia - 1] + Lmp;
      if (j_ri == 2 && c->n_ato_RI == 2) ljj += c->nlmsa[ia - 1];
      abvr[ljj] = abvr[ljj] + charmr;
      if (cal) abvi[ljj] = abvi[ljj] + charmi;
    }
  }
}

/* ---------------------------------------------------------------- mat_solve */
/* Structural pattern of row ii as the reference's symbolic pass marks it,
 * mat_solve_gaussian.f90:85-139 (natome = 0, nspino = 1, regular). */
static void orc_mark_row(const orc_cluster *c, int ii, unsigned char *colnz) {
  const int i = c->newinv[ii];
  if (i > 0) {
    for (int iv = 1; iv <= c->nvois; iv++) {
      int j = IVOIS(c, i, iv);
      int ia = c->numia[j - 1];
      if (ia == 0) {
        colnz[c->newp[j]] = 1;
      } else if (ia > 0) {
        for (int k = c->ianew[ia - 1] + 1; k <= c->ianew[ia - 1] + c->n_ato_RI * c->nlmsa[ia - 1]; k++) colnz[k] = 1;
      } else if (ia == -2 && !c->irregular) {
        for (int k = c->nligneso + 1; k <= c->nligneso + c->nlmso; k++) colnz[k] = 1;
      }
    }
  } else if (i < 0) { /* mat_solve_gaussian.f90:123-137 */
    const int ia = -i;
    for (int k = c->ianew[ia - 1] + 1; k <= c->ianew[ia - 1] + c->n_ato_RI * c->nlmsa[ia - 1]; k++) colnz[k] = 1;
    for (int ib = 1; ib <= c->nbordf[ia - 1]; ib++) colnz[c->newp[IBORD(c, ib, ia)]] = 1;
  } else if (i == 0) {
    for (int k = c->nligneso + 1; k <= c->nligneso + c->nlmso; k++) colnz[k] = 1;
    int lm = ii - c->nligneso;
    int isp = c->iso[lm];
    for (int ib = 1; ib <= c->nsortf; ib++) {
      int j = c->isrt[ib - 1];
      colnz[c->newp[j] - 1 + isp] = 1;
    }
  }
}

typedef void (*mark_fn)(const void *, int, unsigned char *);
typedef void (*fill_fn)(const void *, int, double *, double *, double *, double *);

/* The generic body of mat_solve_gaussian.f90:67-358.  mark() reproduces the
 * symbolic pass per row, fill() stands for the calcMatRow call. */
static int orc_mat_solve_core(int nligne, int nlmso, int cal_comp, const int *lb1, const int *lb2,
                              const void *ud, mark_fn mark, fill_fn fill, double *smr, double *smi,
                              int row_stop, int *fail_row, double *t_fill, double *t_tria, int64_t *nnz_out) {
  unsigned char *colnz = calloc((size_t)nligne + 1, 1);
  int *nb_not_zero = calloc((size_t)nligne + 1, sizeof(int));
  int64_t *lbz = calloc((size_t)nligne + 1, sizeof(int64_t));
  if (!colnz || !nb_not_zero || !lbz) return -1;
  int64_t nb_not_zero_tot = 0;
  int *ind_inv = NULL;
  /* Symbolic passes, :68-152 */
  for (int pass = 1; pass <= 2; pass++) {
    int second = (pass == 2);
    if (second) {
      ind_inv = malloc(sizeof(int) * (size_t)(nb_not_zero_tot > 0 ? nb_not_zero_tot : 1));
      if (!ind_inv) return -1;
    }
    nb_not_zero_tot = 0;
    memset(colnz, 0, (size_t)nligne + 1);
    memset(nb_not_zero, 0, sizeof(int) * ((size_t)nligne + 1));
    for (int ii = nligne; ii >= 1; ii--) {
      mark(ud, ii, colnz);
      for (int j = ii - 1; j >= 1; j--) {
        if (!colnz[j]) continue;
        nb_not_zero_tot++;
        if (second) {
          nb_not_zero[ii]++;
          ind_inv[nb_not_zero_tot - 1] = j;
        }
      }
    }
  }
  int *ind = malloc(sizeof(int) * (size_t)(nb_not_zero_tot > 0 ? nb_not_zero_tot : 1)); /* :154-158 */
  for (int64_t lk = 1; lk <= nb_not_zero_tot; lk++) ind[lk - 1] = ind_inv[nb_not_zero_tot - lk];
  free(ind_inv);
  lbz[1] = 0; /* :162-165 */
  for (int ii = 2; ii <= nligne; ii++) lbz[ii] = lbz[ii - 1] + nb_not_zero[ii - 1];
  if (nnz_out) *nnz_out = nb_not_zero_tot;

  double *abr = calloc((size_t)(nb_not_zero_tot > 0 ? nb_not_zero_tot : 1), sizeof(double));
  double *abi = cal_comp ? calloc((size_t)(nb_not_zero_tot > 0 ? nb_not_zero_tot : 1), sizeof(double)) : NULL;
  double *abvr = calloc((size_t)nligne + 1, sizeof(double));
  double *abvi = cal_comp ? calloc((size_t)nligne + 1, sizeof(double)) : NULL;
  if (!abr || (cal_comp && !abi) || !abvr || (cal_comp && !abvi)) return -1;
  int status = 0;
  double tf = 0, tt = 0;
  if (row_stop < 1) row_stop = 1;
  for (int ii = nligne; ii >= row_stop; ii--) { /* boucle_ii, :196-317 */
    double tp1 = cpu_seconds();
    const int l1 = lb1[ii], l2 = lb2[ii];
    memset(abvr + l1, 0, sizeof(double) * (size_t)(l2 - l1 + 1));
    if (cal_comp) memset(abvi + l1, 0, sizeof(double) * (size_t)(l2 - l1 + 1));
    double *smr_ii = smr + (size_t)(ii - 1) * nlmso;
    double *smi_ii = cal_comp ? smi + (size_t)(ii - 1) * nlmso : NULL;
    fill(ud, ii, abvr, abvi, smr_ii, smi_ii);
    double tp2 = cpu_seconds();
    tf += tp2 - tp1;
    for (int j = l2; j >= ii + 1; j--) { /* :231-262 */
      const double *smr_j = smr + (size_t)(j - 1) * nlmso;
      if (cal_comp) {
        const double facr = abvr[j], faci = abvi[j];
        if (fabs(facr) > EPS15 || fabs(faci) > EPS15) {
          const int64_t b = lbz[j], e = lbz[j] + nb_not_zero[j];
          for (int64_t lk = b; lk < e; lk++) {
            const int k = ind[lk];
            abvr[k] = abvr[k] - facr * abr[lk] + faci * abi[lk];
            abvi[k] = abvi[k] - facr * abi[lk] - faci * abr[lk];
          }
          const double *smi_j = smi + (size_t)(j - 1) * nlmso;
          for (int l = 0; l < nlmso; l++) {
            double sr = smr_ii[l] - facr * smr_j[l] + faci * smi_j[l];
            double si = smi_ii[l] - facr * smi_j[l] - faci * smr_j[l];
            smr_ii[l] = sr;
            smi_ii[l] = si;
          }
        }
      } else {
        const double fac = abvr[j];
        if (fabs(fac) > EPS15) {
          const int64_t b = lbz[j], e = lbz[j] + nb_not_zero[j];
          for (int64_t lk = b; lk < e; lk++) {
            const int k = ind[lk];
            abvr[k] = abvr[k] - fac * abr[lk];
          }
          for (int l = 0; l < nlmso; l++) smr_ii[l] = smr_ii[l] - fac * smr_j[l];
        }
      }
    }
    /* Normalisation, :264-302 */
    if (cal_comp) {
      double den = abvr[ii] * abvr[ii] + abvi[ii] * abvi[ii];
      if (den < 1.e-20) {
        if (fail_row) *fail_row = ii;
        status = 1;
        break;
      }
      double vnr = abvr[ii] / den, vni = -abvi[ii] / den;
      const int64_t b = lbz[ii], e = lbz[ii] + nb_not_zero[ii];
      for (int64_t lk = b; lk < e; lk++) {
        const int k = ind[lk];
        abr[lk] = abvr[k] * vnr - abvi[k] * vni;
        abi[lk] = abvi[k] * vnr + abvr[k] * vni;
      }
      for (int l = 0; l < nlmso; l++) {
        double fr = smr_ii[l] * vnr - smi_ii[l] * vni;
        double fi = smr_ii[l] * vni + smi_ii[l] * vnr;
        smr_ii[l] = fr;
        smi_ii[l] = fi;
      }
    } else {
      if (fabs(abvr[ii]) < 1.e-10) {
        if (fail_row) *fail_row = ii;
        status = 1;
        break;
      }
      double vnr = 1 / abvr[ii];
      const int64_t b = lbz[ii], e = lbz[ii] + nb_not_zero[ii];
      for (int64_t lk = b; lk < e; lk++) abr[lk] = abvr[ind[lk]] * vnr;
      for (int l = 0; l < nlmso; l++) smr_ii[l] = smr_ii[l] * vnr;
    }
    tt += cpu_seconds() - tp2;
  }
  if (status == 0 && row_stop == 1) {
    for (int ii = 2; ii <= nligne; ii++) { /* Resolution, :348-358 */
      double *sr = smr + (size_t)(ii - 1) * nlmso;
      double *si = cal_comp ? smi + (size_t)(ii - 1) * nlmso : NULL;
      for (int64_t lk = lbz[ii]; lk < lbz[ii] + nb_not_zero[ii]; lk++) {
        const int j = ind[lk];
        const double *srj = smr + (size_t)(j - 1) * nlmso;
        if (cal_comp) {
          const double *sij = smi + (size_t)(j - 1) * nlmso;
          for (int l = 0; l < nlmso; l++) {
            double a = sr[l] - abr[lk] * srj[l] + abi[lk] * sij[l];
            double b = si[l] - abr[lk] * sij[l] - abi[lk] * srj[l];
            sr[l] = a;
            si[l] = b;
          }
        } else {
          for (int l = 0; l < nlmso; l++) sr[l] = sr[l] - abr[lk] * srj[l];
        }
      }
    }
  }
  if (t_fill) *t_fill = tf;
  if (t_tria) *t_tria = tt;
  free(colnz); free(nb_not_zero); free(lbz); free(ind); free(abr); free(abi); free(abvr); free(abvi);
  return status;
}

typedef struct {
  row_ctx rc;
  int inc_wave;
  double *scratch_r, *scratch_i;
} grid_ud;

static void grid_mark(const void *ud, int ii, unsigned char *colnz) {
  orc_mark_row(((const grid_ud *)ud)->rc.c, ii, colnz);
}
static void grid_fill(const void *ud, int ii, double *abvr, double *abvi, double *smr_ii, double *smi_ii) {
  const grid_ud *g = (const grid_ud *)ud;
  if (g->inc_wave) {
    orc_calcmatrow(&g->rc, ii, abvr, abvi, smr_ii, smi_ii);
  } else {
    int n = g->rc.c->nlmso;
    memset(g->scratch_r, 0, sizeof(double) * n);
    memset(g->scratch_i, 0, sizeof(double) * n);
    orc_calcmatrow(&g->rc, ii, abvr, abvi, g->scratch_r, g->scratch_i);
  }
}

static int make_ctx(const orc_cluster *c, const double *V, double Enervide, double Eimag, double Ecinetic_out,
                    int E_comp, int cal_comp, int basereel, row_ctx *rc, double **bes, double **neu) {
  /* Cal_comp / Basereel rules of mat.f90:96-104 are the caller's choice;
   * the oracle only requires consistency: E_comp implies cal_comp. */
  if (E_comp && !cal_comp) return -2;
  if (!basereel && !cal_comp) return -2;
  size_t nb = (size_t)c->nsort * (c->lmaxso + 1) * (E_comp ? 2 : 1);
  *bes = malloc(sizeof(double) * (nb ? nb : 1));
  *neu = malloc(sizeof(double) * (nb ? nb : 1));
  orc_phiso(c, E_comp, Ecinetic_out, Eimag, *bes, *neu);
  rc->c = c; rc->V = V; rc->Enervide = Enervide; rc->Eimag = Eimag;
  rc->E_comp = E_comp; rc->cal_comp = cal_comp; rc->basereel = basereel;
  rc->bes = *bes; rc->neu = *neu;
  return 0;
}

int orc_solve(const orc_cluster *c, const double *V, double Enervide, double Eimag, double Ecinetic_out,
              int E_comp, int cal_comp, int basereel, int inc_wave, double *smr, double *smi, int row_stop,
              int *fail_row, double *t_fill, double *t_tria, int64_t *nnz_out) {
  grid_ud g;
  double *bes, *neu;
  int rc = make_ctx(c, V, Enervide, Eimag, Ecinetic_out, E_comp, cal_comp, basereel, &g.rc, &bes, &neu);
  if (rc) return rc;
  g.inc_wave = inc_wave;
  g.scratch_r = calloc(c->nlmso, sizeof(double));
  g.scratch_i = calloc(c->nlmso, sizeof(double));
  rc = orc_mat_solve_core(c->nligne, c->nlmso, cal_comp, c->lb1, c->lb2, &g, grid_mark, grid_fill, smr, smi,
                          row_stop, fail_row, t_fill, t_tria, nnz_out);
  free(g.scratch_r); free(g.scratch_i); free(bes); free(neu);
  return rc;
}

int orc_assemble_dense(const orc_cluster *c, const double *V, double Enervide, double Eimag, double Ecinetic_out,
                       int E_comp, int cal_comp, int basereel, double *a_re, double *a_im, double *rhs_re,
                       double *rhs_im) {
  row_ctx rc;
  double *bes, *neu;
  int st = make_ctx(c, V, Enervide, Eimag, Ecinetic_out, E_comp, cal_comp, basereel, &rc, &bes, &neu);
  if (st) return st;
  const int n = c->nligne, nl = c->nlmso;
  double *br = calloc((size_t)n + 1, sizeof(double)), *bi = calloc((size_t)n + 1, sizeof(double));
  memset(rhs_re, 0, sizeof(double) * (size_t)n * nl);
  if (rhs_im) memset(rhs_im, 0, sizeof(double) * (size_t)n * nl);
  double *dummy = calloc(nl, sizeof(double));
  for (int ii = 1; ii <= n; ii++) {
    memset(br, 0, sizeof(double) * (n + 1));
    memset(bi, 0, sizeof(double) * (n + 1));
    orc_calcmatrow(&rc, ii, br, cal_comp ? bi : NULL, rhs_re + (size_t)(ii - 1) * nl,
                   cal_comp ? rhs_im + (size_t)(ii - 1) * nl : dummy);
    for (int k = 1; k <= n; k++) {
      a_re[(size_t)(k - 1) * n + ii - 1] = br[k];
      if (a_im) a_im[(size_t)(k - 1) * n + ii - 1] = cal_comp ? bi[k] : 0.0;
    }
  }
  free(br); free(bi); free(dummy); free(bes); free(neu);
  return 0;
}

int orc_assemble_csr(const orc_cluster *c, const double *V, double Enervide, double Eimag, double Ecinetic_out,
                     int E_comp, int cal_comp, int basereel, int64_t *row_ptr, int *col, double *val_re,
                     double *val_im, double *rhs_re, double *rhs_im, int64_t *nnz) {
  const int n = c->nligne, nl = c->nlmso;
  unsigned char *mk = calloc((size_t)n + 1, 1);
  /* pattern of row ii: its own diagonal plus the columns the symbolic pass marks */
  int64_t cnt = 0;
  for (int ii = 1; ii <= n; ii++) {
    memset(mk, 0, (size_t)n + 1);
    orc_mark_row(c, ii, mk);
    mk[ii] = 1;
    for (int k = 1; k <= n; k++) cnt += mk[k];
  }
  if (!col) {
    *nnz = cnt;
    free(mk);
    return 0;
  }
  row_ctx rc;
  double *bes, *neu;
  int st = make_ctx(c, V, Enervide, Eimag, Ecinetic_out, E_comp, cal_comp, basereel, &rc, &bes, &neu);
  if (st) { free(mk); return st; }
  double *br = calloc((size_t)n + 1, sizeof(double)), *bi = calloc((size_t)n + 1, sizeof(double));
  double *dummy = calloc(nl, sizeof(double));
  memset(rhs_re, 0, sizeof(double) * (size_t)n * nl);
  if (rhs_im) memset(rhs_im, 0, sizeof(double) * (size_t)n * nl);
  int64_t p = 0;
  row_ptr[0] = 0;
  for (int ii = 1; ii <= n; ii++) {
    memset(mk, 0, (size_t)n + 1);
    orc_mark_row(c, ii, mk);
    mk[ii] = 1;
    const int l1 = c->lb1[ii], l2 = c->lb2[ii];
    for (int k = l1; k <= l2; k++) { br[k] = 0; bi[k] = 0; }
    orc_calcmatrow(&rc, ii, br, cal_comp ? bi : NULL, rhs_re + (size_t)(ii - 1) * nl,
                   cal_comp ? rhs_im + (size_t)(ii - 1) * nl : dummy);
    for (int k = 1; k <= n; k++) {
      if (!mk[k]) continue;
      col[p] = k;
      val_re[p] = br[k];
      if (val_im) val_im[p] = cal_comp ? bi[k] : 0.0;
      p++;
    }
    row_ptr[ii] = p;
  }
  *nnz = p;
  free(mk); free(br); free(bi); free(dummy); free(bes); free(neu);
  return 0;
}

/* ---------------------------------------------------------------- generic */
typedef struct {
  const int64_t *row_ptr;
  const int *col;
  const double *vr, *vi;
  int cal;
} csr_ud;

static void csr_mark(const void *ud, int ii, unsigned char *colnz) {
  const csr_ud *u = (const csr_ud *)ud;
  for (int64_t p = u->row_ptr[ii - 1]; p < u->row_ptr[ii]; p++) colnz[u->col[p]] = 1;
}
static void csr_fill(const void *ud, int ii, double *abvr, double *abvi, double *smr_ii, double *smi_ii) {
  const csr_ud *u = (const csr_ud *)ud;
  (void)smr_ii; (void)smi_ii;
  for (int64_t p = u->row_ptr[ii - 1]; p < u->row_ptr[ii]; p++) {
    abvr[u->col[p]] += u->vr[p];
    if (u->cal) abvi[u->col[p]] += u->vi[p];
  }
}

int orc_mat_solve_csr(int nligne, int nlmso, int cal_comp, const int *lb1, const int *lb2, const int64_t *row_ptr,
                      const int *col, const double *val_re, const double *val_im, double *smr, double *smi,
                      int *fail_row) {
  int *l1 = malloc(sizeof(int) * (nligne + 1)), *l2 = malloc(sizeof(int) * (nligne + 1));
  for (int i = 1; i <= nligne; i++) { l1[i] = lb1[i - 1]; l2[i] = lb2[i - 1]; }
  csr_ud u = {row_ptr, col, val_re, val_im, cal_comp};
  int st = orc_mat_solve_core(nligne, nlmso, cal_comp, l1, l2, &u, csr_mark, csr_fill, smr, smi, 1, fail_row,
                              NULL, NULL, NULL);
  free(l1); free(l2);
  return st;
}

/* ---------------------------------------------------------------- dense */
int orc_dense_solve(int n, int nrhs, double *a_re, double *a_im, double *b_re, double *b_im) {
  double complex *a = malloc(sizeof(double complex) * (size_t)n * n);
  double complex *b = malloc(sizeof(double complex) * (size_t)n * nrhs);
  for (size_t k = 0; k < (size_t)n * n; k++) a[k] = CMPLX(a_re[k], a_im ? a_im[k] : 0.0);
  for (size_t k = 0; k < (size_t)n * nrhs; k++) b[k] = CMPLX(b_re[k], b_im ? b_im[k] : 0.0);
#define A(i, j) a[(size_t)(j) * n + (i)]
#define B(i, j) b[(size_t)(j) * n + (i)]
  int st = 0;
  for (int k = 0; k < n; k++) {
    int p = k;
    double best = cabs(A(k, k));
    for (int i = k + 1; i < n; i++)
      if (cabs(A(i, k)) > best) { best = cabs(A(i, k)); p = i; }
    if (best == 0.0) { st = 1; break; }
    if (p != k) {
      for (int j = 0; j < n; j++) { double complex t = A(k, j); A(k, j) = A(p, j); A(p, j) = t; }
      for (int j = 0; j < nrhs; j++) { double complex t = B(k, j); B(k, j) = B(p, j); B(p, j) = t; }
    }
    double complex piv = A(k, k);
    for (int i = k + 1; i < n; i++) {
      double complex f = A(i, k) / piv;
      if (f == 0) continue;
      for (int j = k + 1; j < n; j++) A(i, j) -= f * A(k, j);
      for (int j = 0; j < nrhs; j++) B(i, j) -= f * B(k, j);
    }
  }
  if (!st) {
    for (int j = 0; j < nrhs; j++)
      for (int i = n - 1; i >= 0; i--) {
        double complex s = B(i, j);
        for (int k = i + 1; k < n; k++) s -= A(i, k) * B(k, j);
        B(i, j) = s / A(i, i);
      }
    for (size_t k = 0; k < (size_t)n * nrhs; k++) {
      b_re[k] = creal(b[k]);
      if (b_im) b_im[k] = cimag(b[k]);
    }
  }
#undef A
#undef B
  free(a); free(b);
  return st;
}

/* ------------------------------------------------------------- invcomp */
/* Restated along the unblocked paths of the reference's vendored LAPACK
 * (sub_util.f): ZGETRF -> ZGETF2, ZTRTRI -> ZTRTI2, and ZGETRI, which
 * invcomp's lwork = n keeps unblocked.  Each helper follows the reference
 * BLAS loop order, operation by operation. */
#define M_(a, n, i, j) (a)[(size_t)((j)-1) * (n) + (i)-1]

/* ONE / z as the reference build evaluates complex division (flang -O2:
 * (a c + b d) / (c^2 + d^2), no scaling); with it the restatement reproduces
 * the compiled reference invcomp bit for bit for n <= 64 */
static double complex orc_crecip(double complex z) {
  double c = creal(z), d = cimag(z), den = c * c + d * d;
  return CMPLX((1.0 * c + 0.0 * d) / den, (0.0 * c - 1.0 * d) / den);
}

/* IZAMAX sub_util.f:23889: first index of the largest |re| + |im| */
static int orc_izamax(int n, const double complex *x) {
  int imax = 1;
  double dmax = fabs(creal(x[0])) + fabs(cimag(x[0]));
  for (int i = 2; i <= n; i++) {
    double d = fabs(creal(x[i - 1])) + fabs(cimag(x[i - 1]));
    if (d > dmax) { imax = i; dmax = d; }
  }
  return imax;
}

/* ZGETF2 sub_util.f:15877-15940 (with ZSWAP, ZSCAL, ZGERU) */
static int orc_zgetf2(int n, double complex *a, int *ipiv) {
  int info = 0;
  const double sfmin = 2.2250738585072014e-308; /* DLAMCH('S') */
  for (int j = 1; j <= n; j++) {
    int jp = j - 1 + orc_izamax(n - j + 1, &M_(a, n, j, j));
    ipiv[j - 1] = jp;
    if (M_(a, n, jp, j) != 0) {
      if (jp != j)
        for (int k = 1; k <= n; k++) {
          double complex t = M_(a, n, j, k);
          M_(a, n, j, k) = M_(a, n, jp, k);
          M_(a, n, jp, k) = t;
        }
      if (j < n) {
        if (cabs(M_(a, n, j, j)) >= sfmin) {
          double complex r = orc_crecip(M_(a, n, j, j));
          for (int i = j + 1; i <= n; i++) M_(a, n, i, j) = r * M_(a, n, i, j);
        } else {
          for (int i = j + 1; i <= n; i++) M_(a, n, i, j) = M_(a, n, i, j) / M_(a, n, j, j);
        }
      }
    } else if (info == 0) {
      info = j;
    }
    if (j < n)
      for (int k = j + 1; k <= n; k++) {
        double complex y = M_(a, n, j, k);
        if (y == 0) continue;
        double complex temp = -1.0 * y;
        for (int i = j + 1; i <= n; i++) M_(a, n, i, k) = M_(a, n, i, k) + M_(a, n, i, j) * temp;
      }
  }
  return info;
}

/* ZTRTRI sub_util.f:15442 (singularity test) -> ZTRTI2 sub_util.f:18626
 * (upper, non-unit) with ZTRMV's column loop and ZSCAL */
static int orc_ztrtri_upper(int n, double complex *a) {
  for (int j = 1; j <= n; j++)
    if (M_(a, n, j, j) == 0) return j;
  for (int j = 1; j <= n; j++) {
    M_(a, n, j, j) = orc_crecip(M_(a, n, j, j));
    double complex ajj = -M_(a, n, j, j);
    double complex *x = &M_(a, n, 1, j);
    for (int jj = 1; jj <= j - 1; jj++) {  /* x := T x, T = inv(U)(1:j-1, 1:j-1) */
      if (x[jj - 1] == 0) continue;
      double complex temp = x[jj - 1];
      for (int i = 1; i <= jj - 1; i++) x[i - 1] = x[i - 1] + temp * M_(a, n, i, jj);
      x[jj - 1] = x[jj - 1] * M_(a, n, jj, jj);
    }
    for (int i = 1; i <= j - 1; i++) x[i - 1] = ajj * x[i - 1];
  }
  return 0;
}

/* ZGETRI sub_util.f:11382-11470, unblocked branch (with ZGEMV and ZSWAP) */
static int orc_zgetri(int n, double complex *a, const int *ipiv) {
  int info = orc_ztrtri_upper(n, a);
  if (info) return info;
  double complex *work = malloc(sizeof(double complex) * (size_t)n);
  for (int j = n; j >= 1; j--) {
    for (int i = j + 1; i <= n; i++) {
      work[i - 1] = M_(a, n, i, j);
      M_(a, n, i, j) = 0;
    }
    if (j < n)  /* A(:, j) += -1 * A(:, j+1:n) work(j+1:n) */
      for (int jj = j + 1; jj <= n; jj++) {
        if (work[jj - 1] == 0) continue;
        double complex temp = -1.0 * work[jj - 1];
        for (int i = 1; i <= n; i++) M_(a, n, i, j) = M_(a, n, i, j) + temp * M_(a, n, i, jj);
      }
  }
  for (int j = n - 1; j >= 1; j--) {
    int jp = ipiv[j - 1];
    if (jp != j)
      for (int i = 1; i <= n; i++) {
        double complex t = M_(a, n, i, j);
        M_(a, n, i, j) = M_(a, n, i, jp);
        M_(a, n, i, jp) = t;
      }
  }
  free(work);
  return 0;
}

/* invcomp mat.f90:1927-1987 with is = 0: ZGETRF, stop on INFO, ZGETRI */
static int orc_invcomp_c(int n, double complex *a) {
  int *ipiv = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int info = orc_zgetf2(n, a, ipiv);
  if (!info) info = orc_zgetri(n, a, ipiv);
  free(ipiv);
  return info;
}

int orc_invcomp(int n, double *a_re, double *a_im) {
  double complex *a = malloc(sizeof(double complex) * (size_t)n * n + 1);
  for (size_t k = 0; k < (size_t)n * n; k++) a[k] = CMPLX(a_re[k], a_im[k]);
  int info = orc_invcomp_c(n, a);
  for (size_t k = 0; k < (size_t)n * n; k++) { a_re[k] = creal(a[k]); a_im[k] = cimag(a[k]); }
  free(a);
  return info;
}

/* ----------------------------------------------------------------- tau */
/* mat.f90:297-405, regular branch (n_RI = 1), nspino = 1 (is1 = is2 =
 * ispinin), Lm01/Lm02 left as the atom row order Lm1/Lm2. */
int orc_tau(int nligne, int nlmso, int nligneso, int cal_comp, int basereel, int fdm_comp,
            const double *x_re, const double *x_im, int natome, const int *ianew, const int *nlmsa,
            double *tau_re, double *tau_im, double *norm_re, double *norm_im) {
  (void)nligne;
  const int n = nlmso;
  const double complex img = CMPLX(0.0, 1.0);
#define SMR(lm, ii) x_re[(size_t)((ii)-1) * n + (lm)-1]
#define SMI(lm, ii) x_im[(size_t)((ii)-1) * n + (lm)-1]
  double complex *norm = malloc(sizeof(double complex) * (size_t)n * n + 1);
  int info = 0;
  if (basereel && n > 0) {
    /* :314-328 (nligneso 1-based = the 0-based first outer row) */
    for (int lm = 1; lm <= n; lm++) {
      int ii = nligneso + lm;
      for (int lm2 = 1; lm2 <= n; lm2++)
        M_(norm, n, lm2, lm) = cal_comp ? CMPLX(-SMI(lm2, ii), SMR(lm2, ii)) : CMPLX(0.0, SMR(lm2, ii));
    }
    for (int lm = 1; lm <= n; lm++) M_(norm, n, lm, lm) = 1.0 + M_(norm, n, lm, lm);
    info = orc_invcomp_c(n, norm);  /* :338 */
    if (norm_re)
      for (size_t k = 0; k < (size_t)n * n; k++) { norm_re[k] = creal(norm[k]); norm_im[k] = cimag(norm[k]); }
    if (info) { free(norm); return info; }
  }
  /* norm_t(k, Lmf) = norm(Lmf, k) (:340) */
  size_t off = 0;
  for (int ia = 0; ia < natome; ia++) {
    int na = nlmsa[ia];
    for (int lm1 = 1; lm1 <= na; lm1++) {
      int ii = ianew[ia] + lm1;
      for (int lm2 = 1; lm2 <= na; lm2++) {
        int jj = ianew[ia] + lm2;
        double complex cfac = 0;
        for (int lmf = 1; lmf <= n; lmf++) {
          double complex a1, a2;
          if (basereel) {
            a1 = 0; a2 = 0;
            for (int k = 1; k <= n; k++) {
              double complex nt = M_(norm, n, lmf, k);
              if (cal_comp) {
                a1 = a1 + nt * CMPLX(SMR(k, ii), SMI(k, ii));
                a2 = a2 + nt * CMPLX(SMR(k, jj), SMI(k, jj));
              } else {
                a1 = a1 + nt * SMR(k, ii);
                a2 = a2 + nt * SMR(k, jj);
              }
            }
          } else {
            a1 = CMPLX(SMR(lmf, ii), cal_comp ? SMI(lmf, ii) : 0.0);
            a2 = CMPLX(SMR(lmf, jj), cal_comp ? SMI(lmf, jj) : 0.0);
          }
          if (fdm_comp) cfac = cfac - img * a1 * a2;
          else cfac = cfac - img * a1 * conj(a2);
        }
        size_t q = off + (size_t)(lm1 - 1) * na + (lm2 - 1);
        tau_re[q] += creal(cfac);
        tau_im[q] += cimag(cfac);
      }
    }
    off += (size_t)na * na;
  }
#undef SMR
#undef SMI
  free(norm);
  return info;
}
#undef M_

/* The symmetrisation of Taull at the end of mat (mat.f90:596-641), one spin,
 * no spin-orbit: with complex harmonics (Ylm_comp) every pair (Lm1, Lm2) and
 * its partner (L2 -m2, L1 -m1) become 0.5 (T + (-1)^(m1+m2) T_partner) and
 * that value times the sign, in the reference's loop order; with real
 * harmonics only i Im(T) is kept.  Not applied with FDM_comp (:600).
 * tau: per atom (nlmsa x nlmsa), row Lm1, harmonics in (L, m) order. */
void orc_tau_sym(int natome, const int *nlmsa, int ylm_comp, double *tau_re, double *tau_im) {
  size_t off = 0;
  for (int ia = 0; ia < natome; ia++) {
    const int na = nlmsa[ia];
    double complex *T = malloc(sizeof(double complex) * (size_t)na * na + 1);
    for (size_t k = 0; k < (size_t)na * na; k++) T[k] = CMPLX(tau_re[off + k], tau_im[off + k]);
#define TT(a, b) T[(size_t)((a)-1) * na + (b)-1]
    int lm1 = 0;
    for (int L1 = 0; lm1 < na; L1++)
      for (int m1 = -L1; m1 <= L1 && lm1 < na; m1++) {
        lm1++;
        int lm2 = 0;
        for (int L2 = 0; lm2 < na; L2++)
          for (int m2 = -L2; m2 <= L2 && lm2 < na; m2++) {
            lm2++;
            if (ylm_comp) {
              const double isg = ((m1 + m2) & 1) ? -1.0 : 1.0;
              const int p2 = lm2 - 2 * m2, p1 = lm1 - 2 * m1;
              const double complex cfac = 0.5 * (TT(lm1, lm2) + isg * TT(p2, p1));
              TT(lm1, lm2) = cfac;
              TT(p2, p1) = isg * cfac;
            } else {
              TT(lm1, lm2) = CMPLX(0.0, cimag(TT(lm1, lm2)));
            }
          }
      }
#undef TT
    for (size_t k = 0; k < (size_t)na * na; k++) {
      tau_re[off + k] = creal(T[k]);
      tau_im[off + k] = cimag(T[k]);
    }
    free(T);
    off += (size_t)na * na;
  }
}

/* Tau_RI of mat.f90:229-295 (complex energy, Test_reg = 0): per atom,
 * Ampl(f, Lm1) = the regular cluster solution at row ianew_reg + Lm1,
 * Ampl_I(f, Lm) = the irregular solution at row ianew_irr + Lm for
 * Lm = 1 .. n_ato_ri * nlmsa (with n_ato_ri = 1 only its first nlmsa
 * columns are set, the second block stays zero);
 * Tau_RI(Lm1, Lm2, i_RI) += sum_f Ampl(f, Lm1) Ampl_I(f, (i_RI-1) nlmsa + Lm2)
 * for i_RI = 1, n_RI.  Solutions (nligne, nlmso) complex, row-major; out per
 * atom (n_RI, nlmsa, nlmsa).  round_regular: the regular solution enters
 * rounded to single precision, as the reference stores it (mat.f90:211). */
void orc_tau_ri(int nlmso, const double *xr_re, const double *xr_im, const int *ianew_reg, const double *xi_re,
                const double *xi_im, const int *ianew_irr, int natome, const int *nlmsa, int n_ato_ri, int n_ri,
                int round_regular, double *out_re, double *out_im) {
  size_t off = 0;
  for (int ia = 0; ia < natome; ia++) {
    const int na = nlmsa[ia];
    for (int lm1 = 1; lm1 <= na; lm1++) {
      const size_t ii = (size_t)(ianew_reg[ia] + lm1 - 1) * nlmso;
      int lm = 0;
      for (int i_ri = 1; i_ri <= n_ri; i_ri++)
        for (int lm2 = 1; lm2 <= na; lm2++) {
          lm++;
          const size_t q = off + ((size_t)(i_ri - 1) * na + (lm1 - 1)) * na + (lm2 - 1);
          /* each product goes straight into Tau_RI, as :286-288 adds it */
          double complex t = CMPLX(out_re[q], out_im[q]);
          if (lm <= n_ato_ri * na) {
            const size_t jj = (size_t)(ianew_irr[ia] + lm - 1) * nlmso;
            for (int f = 0; f < nlmso; f++) {
              /* mat.f90:211 keeps the regular solution as cmplx(smr, smi)
               * of default kind: each part rounded to single precision */
              const double ar = round_regular ? (double)(float)xr_re[ii + f] : xr_re[ii + f];
              const double ai = round_regular ? (double)(float)xr_im[ii + f] : xr_im[ii + f];
              const double br = xi_re[jj + f], bi = xi_im[jj + f];
              t = CMPLX(creal(t) + (ar * br - ai * bi), cimag(t) + (ar * bi + ai * br));
            }
          }
          out_re[q] = creal(t);
          out_im[q] = cimag(t);
        }
    }
    off += (size_t)n_ri * na * na;
  }
}

/* ---------------------------------------------------------------- Sphere */
/* Cal_Hankel (Second_kind = F) and Cal_Bessel, sphere.f90:7255-7350: the
 * normalised h_L(k r), j_L(k r) and their derivatives in r, L = 0..lmax */
static void orc_hankel_bessel(double complex konde, int lmax, double r, double complex *hs, double complex *dhs,
                              double complex *bs, double complex *dbs) {
  double complex z = konde * r;
  double complex ht[64], bt[64];
  ht[0] = -I;
  ht[1] = -1.0 - I / z;
  for (int L = 2; L <= lmax + 1; L++) ht[L] = ((double)(2 * L - 1) * ht[L - 1]) / z - ht[L - 2];
  double complex fnorm = csqrt(konde / PI);
  double complex cfac = fnorm * cexp(I * z) / z;
  for (int L = 0; L <= lmax; L++) {
    double complex d = ((double)L * ht[L]) / z - ht[L + 1];
    hs[L] = cfac * ht[L];
    dhs[L] = cfac * konde * d;
  }
  bt[0] = csin(z);
  bt[1] = -ccos(z) + bt[0] / z;
  for (int L = 2; L <= lmax + 1; L++) bt[L] = ((double)(2 * L - 1) * bt[L - 1]) / z - bt[L - 2];
  cfac = fnorm / z;
  for (int L = 0; L <= lmax; L++) {
    double complex d = ((double)L * bt[L]) / z - bt[L + 1];
    bs[L] = cfac * bt[L];
    dbs[L] = cfac * konde * d;
  }
}

/* f_interp2, tools.f90:504-522 */
static double orc_interp2(double r, double rm, double r0, double rp, double fm, double f0, double fp) {
  double dp0 = rp - r0, dpm = rp - rm, d0m = r0 - rm;
  double a = (fm * dp0 - f0 * dpm + fp * d0m) / (d0m * dp0 * dpm);
  double b = (f0 - fm) / d0m - a * (r0 + rm);
  double c = f0 - a * (r0 * r0) - b * r0;
  return a * (r * r) + b * r + c;
}

/* Sphere's work with, when u_re is given, the renormalised radial solutions
 * of every L (u_re/u_im[L * nr + ir]) kept as well. */
static int sphere_impl(int nr, const double *r, const double *Vr, int numat, double rmtg, int lmax, int relativiste,
                       double Ecinetic, double Eimag, double Enervide, double V_intmax, int nbord, const double *bxyz,
                       const double *pos, const double *ylmato, int nlm, double *phi_re, double *phi_im,
                       double *tau_re, double *tau_im, double *u_re, double *u_im, double *phi2_re,
                       double *phi2_im) {
  if (nr < 4 || lmax < 0 || lmax > 60 || nlm < (lmax + 1) * (lmax + 1)) return -1;
  const int n = nr;
  double *V = malloc(sizeof(double) * n), *f2 = malloc(sizeof(double) * n), *g0 = malloc(sizeof(double) * n),
         *gm = malloc(sizeof(double) * n), *gp = malloc(sizeof(double) * n), *gpi = malloc(sizeof(double) * n),
         *cgm = malloc(sizeof(double) * n), *cgp = malloc(sizeof(double) * n), *cg0 = malloc(sizeof(double) * n),
         *ur = malloc(sizeof(double) * n), *ui = malloc(sizeof(double) * n);
  if (!V || !f2 || !g0 || !gm || !gp || !gpi || !cgm || !cgp || !cg0 || !ur || !ui) return -1;
  /* mod_V, sphere.f90:216-268: spherical potential, capped at V_intmax */
  for (int ir = 0; ir < n; ir++) V[ir] = Vr[ir];
  if (V_intmax < 1000.0)
    for (int ir = 0; ir < n; ir++) V[ir] = fmin(V[ir], V_intmax);
  int nrmtg;
  for (nrmtg = 1; nrmtg <= n - 1; nrmtg++)
    if (r[nrmtg - 1] > rmtg + EPS10) break;
  /* coef_sch_rad, sphere.f90:324-405 */
  for (int i = 1; i <= n; i++) {
    double rp = (i == n) ? r[n - 1] * r[n - 1] / r[n - 2] : r[i];
    double rm = (i == 1) ? r[0] * r[0] / r[1] : r[i - 2];
    double r0 = r[i - 1];
    double dr = 0.5 * (rp - rm);
    double claplm = 1 / ((r0 - rm) * dr), claplp = 1 / ((rp - r0) * dr);
    double clapl0 = -claplm - claplp;
    if (relativiste) {
      cgm[i - 1] = (rp - r0) / ((rm - r0) * (rp - rm));
      cgp[i - 1] = (rm - r0) / ((rp - r0) * (rm - rp));
      cg0[i - 1] = -cgm[i - 1] - cgp[i - 1];
    }
    f2[i - 1] = 1 / (r[i - 1] * r[i - 1]);
    double vme = V[i - 1] - Enervide;
    g0[i - 1] = -clapl0 + vme;
    gm[i - 1] = -claplm;
    gp[i - 1] = -claplp;
    if (relativiste) {
      const double a2s4 = 0.25 * (ALFA_SF * ALFA_SF);
      double bder = 1 / (1 - a2s4 * vme), dvr;
      if (i == 1)
        dvr = (double)(2 * numat) / (r[0] * r[0]);
      else if (i == n)
        dvr = cgm[i - 1] * V[i - 2] + cg0[i - 1] * V[i - 1] + cgp[i - 1] * V[i - 1];
      else
        dvr = cgm[i - 1] * V[i - 2] + cg0[i - 1] * V[i - 1] + cgp[i - 1] * V[i];
      double fac = a2s4 * bder * dvr;
      g0[i - 1] = g0[i - 1] - a2s4 * (vme * vme) - fac * (cg0[i - 1] - 1 / r[i - 1]);
      gm[i - 1] = gm[i - 1] - fac * cgm[i - 1];
      gp[i - 1] = gp[i - 1] - fac * cgp[i - 1];
    }
  }
  for (int i = 0; i < n - 1; i++) gpi[i] = 1 / gp[i];
  const int ecomp = fabs(Eimag) > EPS10 || Ecinetic < EPS10; /* Radial_comp = Ecomp (no Hubbard) */
  const double complex konde = csqrt(CMPLX(Ecinetic, Eimag));
  double complex hs[64], dhs[64], bs[64], dbs[64];
  orc_hankel_bessel(konde, lmax, rmtg, hs, dhs, bs, dbs);
  const int inr = nrmtg < n ? nrmtg : n;
  const double d12 = r[inr - 2] - r[inr - 3], d01 = r[inr - 1] - r[inr - 2], d02 = r[inr - 1] - r[inr - 3];
  const double d2 = 1 / (d12 * d02), d1 = 1 / (d12 * d01), d0 = 1 / (d01 * d02);
  double *rrel = malloc(sizeof(double) * (nbord > 0 ? nbord : 1));
  int *iint = malloc(sizeof(int) * (nbord > 0 ? nbord : 1));
  for (int ib = 0; ib < nbord; ib++) { /* posrel (P1): distance to the atom; cal_phiato's three radii */
    double u0 = bxyz[3 * ib] - pos[0], u1 = bxyz[3 * ib + 1] - pos[1], u2 = bxyz[3 * ib + 2] - pos[2];
    rrel[ib] = sqrt(u0 * u0 + u1 * u1 + u2 * u2);
    int i;
    for (i = 2; i <= n; i++)
      if (r[i - 1] > rrel[ib]) break;
    if (i > n) i = n;
    if (i < 3) i = 3;
    iint[ib] = i;
  }
  if (nbord > 0) {
    memset(phi_re, 0, sizeof(double) * (size_t)nbord * nlm);
    memset(phi_im, 0, sizeof(double) * (size_t)nbord * nlm);
    if (phi2_re) {
      memset(phi2_re, 0, sizeof(double) * (size_t)nbord * nlm);
      memset(phi2_im, 0, sizeof(double) * (size_t)nbord * nlm);
    }
  }
  if (phi2_re && nrmtg > n - 1) return -1; /* the irregular solution starts one point beyond the matching point */
  double complex *us = phi2_re ? malloc(sizeof(double complex) * n) : NULL;
  double *hb = malloc(sizeof(double) * 2 * (lmax + 1)), *hn = malloc(sizeof(double) * 2 * (lmax + 1));
  for (int L = 0; L <= lmax; L++) {
    /* Sch_radial, sphere.f90:413-730 (one spin, no spin-orbit, no Hubbard) */
    const int L2 = L * (L + 1);
    const double ci = -Eimag / (4 * L + 6);
    const double br = -(numat / (L + 1.0));
    const double cr = -((double)(2 * numat) * br + Ecinetic) / (4 * L + 6);
    double p = relativiste ? sqrt((double)(L * L + L + 1) - (ALFA_SF * numat) * (ALFA_SF * numat)) : L + 1.0;
    for (int k = 0; k < 2; k++) {
      ur[k] = 1.0 * (1 + br * r[k] + cr * (r[k] * r[k])) * pow(r[k], p);
      ui[k] = ecomp ? 1.0 * ci * pow(r[k], p + 2) : 0.0;
    }
    for (int i = 2; i <= n - 1; i++) { /* 1-based ir; im = ir - 1, ip = ir + 1 */
      const int ir = i - 1, im = i - 2, ip = i;
      double td = g0[ir] + L2 * f2[ir];
      double a = td * ur[ir] + gm[ir] * ur[im];
      double b = 0;
      if (ecomp) {
        b = td * ui[ir] + gm[ir] * ui[im];
        a = a + Eimag * ui[ir];
        b = b - Eimag * ur[ir];
      }
      ur[ip] = -a * gpi[ir];
      ui[ip] = ecomp ? -b * gpi[ir] : 0.0;
    }
    for (int i = 0; i < n; i++) {
      ur[i] = ur[i] / r[i];
      if (ecomp) ui[i] = ui[i] / r[i];
    }
    /* Renormal, sphere.f90:931-1150, and cal_ampl :1157-1190 (nlm1 = nlm2 = 1, nspino = 1) */
    double a = ur[inr - 3] * d2 - ur[inr - 2] * d1 + ur[inr - 1] * d0;
    double b = (ur[inr - 2] - ur[inr - 3]) / d12 - a * (r[inr - 2] + r[inr - 3]);
    double c = ur[inr - 1] - a * (r[inr - 1] * r[inr - 1]) - b * r[inr - 1];
    double ai = 0, bi = 0, cc = 0;
    if (ecomp) {
      ai = ui[inr - 3] * d2 - ui[inr - 2] * d1 + ui[inr - 1] * d0;
      bi = (ui[inr - 2] - ui[inr - 3]) / d12 - ai * (r[inr - 2] + r[inr - 3]);
      cc = ui[inr - 1] - ai * (r[inr - 1] * r[inr - 1]) - bi * r[inr - 1];
    }
    double complex u1 = CMPLX(a * (rmtg * rmtg) + b * rmtg + c, ai * (rmtg * rmtg) + bi * rmtg + cc);
    double complex u2 = CMPLX(2 * a * rmtg + b, 2 * ai * rmtg + bi);
    double wout = 1 / (PI * (rmtg * rmtg));
    double complex s1 = -I * hs[L], s2 = -I * dhs[L];
    double complex ws = u1 * s2 - u2 * s1, we = u1 * dbs[L] - u2 * bs[L];
    double complex ampl = wout / we;
    double complex tau = -we / ws;
    if (tau_re) tau_re[L] = creal(tau);
    if (tau_im) tau_im[L] = cimag(tau);
    for (int i = 0; i < n; i++) {
      double complex z = ampl * (ecomp ? CMPLX(ur[i], ui[i]) : CMPLX(ur[i], 0.0));
      ur[i] = creal(z);
      ui[i] = ecomp ? cimag(z) : 0.0;
    }
    if (u_re)
      for (int i = 0; i < n; i++) {
        u_re[(size_t)L * n + i] = ur[i];
        u_im[(size_t)L * n + i] = ui[i];
      }
    /* the irregular solution (Sch_radial_solsing, sphere.f90:2591-2806, Classic_irreg, as Sphere calls it
     * at :155-162): -i r h_L(k r) from the matching point outwards (cbessneu / cbessneur), integrated
     * inwards by the same three-point recurrence with the roles of gm and gp exchanged, then over r */
    if (us) {
      for (int ir = nrmtg; ir <= n; ir++) {
        double complex h;
        if (ecomp) {
          const double complex fnc = csqrt(konde / PI), z = konde * r[ir - 1];
          orc_cbessneu(creal(fnc), cimag(fnc), creal(z), cimag(z), lmax, hb, hn);
          h = CMPLX(hb[2 * L], hb[2 * L + 1]) + I * CMPLX(hn[2 * L], hn[2 * L + 1]);
        } else {
          const double kr = creal(konde), fn = sqrt(kr / PI);
          orc_cbessneur(fn, kr * r[ir - 1], lmax, hb, hn);
          h = hb[L] + I * hn[L];
        }
        us[ir - 1] = -I * r[ir - 1] * h;
      }
      for (int ir = nrmtg; ir >= 2; ir--) {
        const int ip = ir - 1, im = ir + 1;
        const double td = g0[ir - 1] + L2 * f2[ir - 1];
        double complex v = td * us[ir - 1] + gp[ir - 1] * us[im - 1];
        if (ecomp) v = v - I * Eimag * us[ir - 1];
        us[ip - 1] = -v * (1 / gm[ir - 1]);
      }
    }
    /* cal_phiato, sphere.f90:3186-3345 */
    const int Lm0 = L * L + L + 1;
    for (int ib = 0; ib < nbord; ib++) {
      const int i = iint[ib];
      const double rm = r[i - 3], r0 = r[i - 2], rp = r[i - 1];
      double ph = orc_interp2(rrel[ib], rm, r0, rp, ur[i - 3], ur[i - 2], ur[i - 1]);
      double phi = ecomp ? orc_interp2(rrel[ib], rm, r0, rp, ui[i - 3], ui[i - 2], ui[i - 1]) : 0.0;
      for (int mpp = -L; mpp <= L; mpp++) {
        const int Lm = Lm0 + mpp; /* 1-based */
        const double y = ylmato[(size_t)ib * nlm + Lm - 1];
        phi_re[(size_t)ib * nlm + Lm - 1] = phi_re[(size_t)ib * nlm + Lm - 1] + ph * y;
        if (ecomp) phi_im[(size_t)ib * nlm + Lm - 1] = phi_im[(size_t)ib * nlm + Lm - 1] + phi * y;
      }
      if (us) { /* i_RI = 2: the irregular solution over r (its imaginary part only with Radial_comp) */
        const double sm = creal(us[i - 3]) / r[i - 3], s0 = creal(us[i - 2]) / r[i - 2], sp = creal(us[i - 1]) / r[i - 1];
        const double p2 = orc_interp2(rrel[ib], rm, r0, rp, sm, s0, sp);
        double p2i = 0;
        if (ecomp)
          p2i = orc_interp2(rrel[ib], rm, r0, rp, cimag(us[i - 3]) / r[i - 3], cimag(us[i - 2]) / r[i - 2],
                            cimag(us[i - 1]) / r[i - 1]);
        for (int mpp = -L; mpp <= L; mpp++) {
          const size_t o = (size_t)ib * nlm + Lm0 + mpp - 1;
          const double y = ylmato[o];
          phi2_re[o] = phi2_re[o] + p2 * y;
          if (ecomp) phi2_im[o] = phi2_im[o] + p2i * y;
        }
      }
    }
  }
  free(us); free(hb); free(hn);
  free(V); free(f2); free(g0); free(gm); free(gp); free(gpi); free(cgm); free(cgp); free(cg0); free(ur); free(ui);
  free(rrel); free(iint);
  return ecomp;
}

int orc_sphere(int nr, const double *r, const double *Vr, int numat, double rmtg, int lmax, int relativiste,
               double Ecinetic, double Eimag, double Enervide, double V_intmax, int nbord, const double *bxyz,
               const double *pos, const double *ylmato, int nlm, double *phi_re, double *phi_im, double *tau_re,
               double *tau_im) {
  return sphere_impl(nr, r, Vr, numat, rmtg, lmax, relativiste, Ecinetic, Eimag, Enervide, V_intmax, nbord, bxyz, pos,
                     ylmato, nlm, phi_re, phi_im, tau_re, tau_im, NULL, NULL, NULL, NULL);
}

int orc_sphere_ri(int nr, const double *r, const double *Vr, int numat, double rmtg, int lmax, int relativiste,
                  double Ecinetic, double Eimag, double Enervide, double V_intmax, int nbord, const double *bxyz,
                  const double *pos, const double *ylmato, int nlm, double *phi_re, double *phi_im, double *phi2_re,
                  double *phi2_im, double *tau_re, double *tau_im) {
  if (!phi2_re || !phi2_im) return -1;
  return sphere_impl(nr, r, Vr, numat, rmtg, lmax, relativiste, Ecinetic, Eimag, Enervide, V_intmax, nbord, bxyz, pos,
                     ylmato, nlm, phi_re, phi_im, tau_re, tau_im, NULL, NULL, phi2_re, phi2_im);
}

/* ------------------------------------------------------------------------
 * The cross-section tensors of the absorbing atom (tensor.f90, sphere.f90
 * Radial_integral): radial integrals of the golden rule, Gaunt
 * coefficients, tens_ab's contraction of rof x rof x Taull and tenseur_car's
 * Cartesian tensors.
 * ------------------------------------------------------------------------ */

/* f_integr3, tools.f90:308-419, with ir0 = 1: the integral of fct from 0 to
 * Radius, piecewise cubic through four grid points (the first interval from
 * 0 with the value at 0 extrapolated by a parabola; the last interval, cut at
 * Radius, reuses the cubic of the interval before it). */
double orc_f_integr3(const double *r, const double *fct, int nrm, double Radius) {
#define RR(i) r[(i)-1]
#define FF(i) fct[(i)-1]
  double res = 0;
  if (Radius < EPS10) return 0;
  if (nrm < 4) return NAN;
  const double tiers = 1.0 / 3;
  int the_end = 0;
  double r1 = 0, r2 = 0, r3 = 0, r4 = 0, f1 = 0, f2 = 0, f3 = 0, f4 = 0, rm = 0, rp = 0;
  for (int i = 1; i <= nrm - 1; i++) {
    if (i == 1) {
      const double rmm = RR(1), r0 = RR(2), rpp = RR(3), fm = FF(1), f0 = FF(2), fp = FF(3);
      if (fabs(fp) < 1e-20 && fabs(f0) < 1e-20 && fabs(fm) < 1e-20) continue;
      const double dp0 = rpp - r0, dpm = rpp - rmm, d0m = r0 - rmm;
      const double a = (fm * dp0 - f0 * dpm + fp * d0m) / (d0m * dp0 * dpm);
      const double b = (f0 - fm) / d0m - a * (r0 + rmm);
      const double c = f0 - a * (r0 * r0) - b * r0;
      f1 = c;
      r1 = 0;
      r2 = RR(i); f2 = FF(i);
      r3 = RR(i + 1); f3 = FF(i + 1);
      r4 = RR(i + 2); f4 = FF(i + 2);
      rm = 0;
      rp = r3;
    } else if (i == nrm - 1 || RR(i + 1) > Radius) {
      rm = RR(i);
      rp = RR(nrm);
      if (i < 3) {
        r1 = RR(i - 1); f1 = FF(i - 1);
        r2 = RR(i); f2 = FF(i);
        r3 = RR(i + 1); f3 = FF(i + 1);
        r4 = RR(i + 2); f4 = FF(i + 2);
      }
    } else {
      r1 = RR(i - 1); f1 = FF(i - 1);
      r2 = RR(i); f2 = FF(i);
      r3 = RR(i + 1); f3 = FF(i + 1);
      r4 = RR(i + 2); f4 = FF(i + 2);
      rm = r2;
      rp = r3;
    }
    if (rp > Radius) {
      rp = Radius;
      the_end = 1;
    }
    if (fabs(f2) < 1e-20 && fabs(f3) < 1e-20) continue;
    const double rap14 = (f1 - f4) / (r1 - r4), rap24 = (f2 - f4) / (r2 - r4), rap34 = (f3 - f4) / (r3 - r4);
    const double fac1 = (rap14 - rap34) / (r1 - r3), fac2 = (rap24 - rap34) / (r2 - r3);
    const double a = (fac1 - fac2) / (r1 - r2);
    const double b = fac1 - a * (r1 + r3 + r4);
    const double c = rap14 - a * (r1 * r1 + r1 * r4 + r4 * r4) - b * (r1 + r4);
    const double d = f1 - (a * (r1 * r1) + b * r1 + c) * r1;
    res = res + (0.25 * (rm + rp) * (rm * rm + rp * rp) * a + tiers * (rm * rm + rm * rp + rp * rp) * b +
                 0.5 * (rm + rp) * c + d) * (rp - rm);
    if (the_end) break;
  }
  return res;
#undef RR
#undef FF
}

/* coef_init, general.f90:774-839: the core states |j, mj> of the edge(s),
 * their spin-up / spin-down weights coef_g[init][isp] and orbital m_g. */
int orc_coef_init(int jseuil, int lseuil, int nbseuil, int ninit, double *coef_g, int *m_g, int *is_g) {
  for (int k = 0; k < 2 * ninit; k++) {
    coef_g[k] = 0;
    m_g[k] = 1000;
  }
  int is_init = (jseuil == 1 || jseuil == 3 || jseuil == 5 || jseuil == 7) ? 1 : -1;
  int init = 0;
  for (int iseuil = 1; iseuil <= nbseuil; iseuil++) {
    if (iseuil == 2) is_init = -is_init;
    const int li = lseuil;
    const double J = li + 0.5 * is_init;
    const int ni = (int)lround(2 * J + 1);
    for (int i = 1; i <= ni; i++) {
      if (init >= ninit) return -1;
      init++;
      const double Jz = -J + i - 1;
      is_g[init - 1] = is_init;
      for (int isp = 1; isp <= 2; isp++) {
        const double spin = isp == 1 ? 0.5 : -0.5;
        const int is = isp == 1 ? is_init : 1;
        const int mi = (int)lround(Jz - spin);
        if (abs(mi) > li) continue;
        const double ci2 = (li + 0.5 + 2 * spin * is_init * Jz) / (2 * li + 1);
        if (ci2 < EPS6) continue;
        coef_g[2 * (init - 1) + isp - 1] = is * sqrt(ci2);
        m_g[2 * (init - 1) + isp - 1] = mi;
      }
    }
  }
  return init;
}

static double fact_d(int n) {
  double f = 1;
  for (int k = 2; k <= n; k++) f *= k;
  return f;
}

/* <l1 m1 l2 m2 | l3 m1+m2>, the Clebsch-Gordan coefficient that cgc
 * (tensor.f90:1678-1790, Rose eq. 3.18) computes, here by Racah's closed
 * formula (Edmonds, Angular Momentum in Quantum Mechanics, eq. 3.6.11). */
double orc_cgc(int l1, int l2, int l3, int m1, int m2) {
  const int m3 = m1 + m2;
  if (l3 < abs(l1 - l2) || l3 > l1 + l2 || abs(m1) > l1 || abs(m2) > l2 || abs(m3) > l3) return 0;
  const double pre = sqrt((2 * l3 + 1) * fact_d(l1 + l2 - l3) * fact_d(l1 - l2 + l3) * fact_d(-l1 + l2 + l3) /
                          fact_d(l1 + l2 + l3 + 1) * fact_d(l1 + m1) * fact_d(l1 - m1) * fact_d(l2 + m2) *
                          fact_d(l2 - m2) * fact_d(l3 + m3) * fact_d(l3 - m3));
  double s = 0;
  for (int k = 0; k <= l1 + l2 - l3; k++) {
    const int a = l1 + l2 - l3 - k, b = l1 - m1 - k, c = l2 + m2 - k, d = l3 - l2 + m1 + k, e = l3 - l1 - m2 + k;
    if (a < 0 || b < 0 || c < 0 || d < 0 || e < 0) continue;
    s += ((k & 1) ? -1.0 : 1.0) / (fact_d(k) * fact_d(a) * fact_d(b) * fact_d(c) * fact_d(d) * fact_d(e));
  }
  return pre * s;
}

static int ipow_m1(int k) { return (k & 1) ? -1 : 1; }

/* Gaunt_r, tensor.f90:1796-1866: the integral of three real harmonics
 * Y_l1m1 Y_l2m2 Y_l3m3, from complex-harmonic Clebsch-Gordan products. */
double orc_gaunt_r(int l1, int m1, int l2, int m2, int l3, int m3) {
  if ((l3 + l1 + l2) % 2 != 0 || abs(l1 - l3) > l2 || (l1 + l3) < l2) return 0;
  const int mi = abs(m1), mgam = abs(m2), m = abs(m3);
  if (mi + mgam != m && mi + m != mgam && m + mgam != mi) return 0;
  const double sqr2i = 1 / sqrt(2.0);
  const int ltsti = l1, itr = l2, ltst = l3;
  const double pre =
      sqrt((2 * ltst + 1) * (2 * itr + 1) / ((2 * ltsti + 1) * (4 * PI))) * orc_cgc(ltst, itr, ltsti, 0, 0);
  const int itypi = m1 == 0 ? 0 : mi / m1, itygam = m2 == 0 ? 0 : mgam / m2, ityp = m3 == 0 ? 0 : m / m3;
  const int nexp = abs(itypi) + abs(itygam) + abs(ityp);
  int ns = itypi + ityp + itygam;
  if (ns == 0) ns = 1;
  const int nsf = ns > 0 ? 1 : -1;
  double sp = 1;
  for (int k = 0; k < nexp; k++) sp *= sqr2i;
  const double gf = nsf * ipow_m1(mi) * sp * pre;
  double ab;
  if (m + mgam + mi == 0) {
    ab = orc_cgc(ltst, itr, ltsti, 0, 0);
  } else {
    const int f1 = itypi + itygam * ityp, f2 = itygam + itypi * ityp, f3 = ityp + itypi * itygam;
    ab = 0;
    if (m + mgam == mi && f1 != 0) ab = ipow_m1(mi) * orc_cgc(ltst, itr, ltsti, m, mgam) * f1;
    if (m + mi == mgam && f2 != 0) ab = ab + ipow_m1(mgam) * orc_cgc(ltst, itr, ltsti, -m, mgam) * f2;
    if (mgam + mi == m && f3 != 0) ab = ab + ipow_m1(m) * orc_cgc(ltst, itr, ltsti, m, -mgam) * f3;
  }
  return gf * ab;
}

/* Gaunt_xrc, tensor.f90:1870-1912, real harmonics (Ylm_comp false): Y_lm
 * and Y_lomo real, Y_limi complex. */
void orc_gaunt_xrc(int l, int m, int lo, int mo, int li, int mi, double *gr, double *gi) {
  const double s2 = sqrt(2.0);
  if (mi == 0) {
    *gr = orc_gaunt_r(l, m, lo, mo, li, mi);
    *gi = 0;
  } else if (mi > 0) {
    *gr = ipow_m1(mi) * orc_gaunt_r(l, m, lo, mo, li, mi) / s2;
    *gi = ipow_m1(mi) * orc_gaunt_r(l, m, lo, mo, li, -mi) / s2;
  } else {
    *gr = orc_gaunt_r(l, m, lo, mo, li, -mi) / s2;
    *gi = -orc_gaunt_r(l, m, lo, mo, li, mi) / s2;
  }
}

/* Radial_integral, sphere.f90:1490-1680, and radial_matrix, :2209-2310 (one
 * spin, spherical potential, no spin-orbit, Hubbard or full potential, no
 * TDDFT or NRIXS): Sphere's renormalised regular solution u_L (Sch_radial),
 * then rof(L, ip, is) = fac(ip) * integral_0^Rmtsd psii_is(r) r u_L(r) r^ip dr
 * (f_integr3) with fac 0.5 alfa_sf, 1, Vecond / 2, -Vecond^2 / 6 for ip = 0..3,
 * Vecond = alfa_sf / 2 * max(0.001, Energ + Eseuil(is)); the imaginary part
 * from Im u_L.  psii[is * nr + ir]; rof out [((is * nip) + ip - ip0) *
 * (lmax + 1) + L].  Returns Radial_comp (1 complex solutions, 0 real) or -1. */
int orc_radial_integral(int nr, const double *r, const double *Vr, int numat, double rmtg, double rmtsd, int lmax,
                        int relativiste, double Ecinetic, double Eimag, double Enervide, double Energ,
                        double V_intmax, int nbseuil, const double *Eseuil, const double *psii, int ip0, int ip_max,
                        double *rof_re, double *rof_im) {
  if (ip0 < 0 || ip_max > 3 || ip_max < ip0 || nbseuil < 1 || nbseuil > 2) return -1;
  double *ur = malloc(sizeof(double) * (size_t)nr * (lmax + 1)), *ui = malloc(sizeof(double) * (size_t)nr * (lmax + 1));
  double *fct = malloc(sizeof(double) * nr), *rb = malloc(sizeof(double) * nr);
  const int ecomp = sphere_impl(nr, r, Vr, numat, rmtg, lmax, relativiste, Ecinetic, Eimag, Enervide, V_intmax, 0,
                                NULL, NULL, NULL, (lmax + 1) * (lmax + 1), NULL, NULL, NULL, NULL, ur, ui, NULL, NULL);
  if (ecomp < 0) return -1;
  const int nip = ip_max - ip0 + 1;
  double vecond[2];
  for (int i = 0; i < nbseuil; i++) vecond[i] = 0.5 * ALFA_SF * fmax(0.001, Energ + Eseuil[i]);
  for (int L = 0; L <= lmax; L++)
    for (int ip = ip0; ip <= ip_max; ip++) {
      for (int ir = 0; ir < nr; ir++) rb[ir] = ip == 0 ? 1.0 : ip == 1 ? r[ir] : ip == 2 ? r[ir] * r[ir]
                                                                                    : r[ir] * r[ir] * r[ir];
      for (int is = 0; is < nbseuil; is++) {
        const double fac = ip == 0 ? 0.5 * ALFA_SF : ip == 1 ? 1.0 : ip == 2 ? 0.5 * vecond[is]
                                                                          : -(1.0 / 6) * (vecond[is] * vecond[is]);
        for (int ir = 0; ir < nr; ir++) fct[ir] = psii[(size_t)is * nr + ir] * r[ir] * ur[(size_t)L * nr + ir] * rb[ir];
        const double radlr = fac * orc_f_integr3(r, fct, nr, rmtsd);
        double radli = 0;
        if (ecomp) {
          for (int ir = 0; ir < nr; ir++)
            fct[ir] = psii[(size_t)is * nr + ir] * r[ir] * ui[(size_t)L * nr + ir] * rb[ir];
          radli = fac * orc_f_integr3(r, fct, nr, rmtsd);
        }
        const size_t o = ((size_t)is * nip + ip - ip0) * (lmax + 1) + L;
        rof_re[o] = radlr;
        rof_im[o] = radli;
      }
    }
  free(ur); free(ui); free(fct); free(rb);
  return ecomp;
}

/* tens_ab, tensor.f90:947-1345, for the branch of an FDM absorption
 * calculation: no TDDFT, NRIXS, Green's-function or FDM_comp output, no
 * singular solution, no spin-orbit, one final-state spin, spherical
 * potential, real harmonics, electric multipoles (irang, jrang >= 1). */
static void tens_ab_r(const orc_tens_in *in, int ninitr, int irang, int jrang, int l_s, int m_s, int l_i, int m_i,
                      const double complex *rof, const double complex *taull, double complex *Ten) {
  const int li = in->lseuil, nlm = in->nlm_probe, nip = in->ip_max - in->ip0 + 1;
  for (int k = 0; k < ninitr; k++) Ten[k] = 0;
  for (int ig = 1; ig <= in->ninit; ig++) {
    const int iseuil = ig <= in->ninit1 ? 1 : 2;
    const int initr = in->core_resolved ? ig : iseuil;
    for (int isg1 = 1; isg1 <= 2; isg1++) {
      const int mi1 = in->m_g[2 * (ig - 1) + isg1 - 1];
      const double ci1 = in->coef_g[2 * (ig - 1) + isg1 - 1];
      if (fabs(ci1) < EPS6) continue;
      for (int isg2 = 1; isg2 <= 2; isg2++) {
        if (li == 0 && isg1 != isg2) continue;
        const int mi2 = in->m_g[2 * (ig - 1) + isg2 - 1];
        const double ci2 = in->coef_g[2 * (ig - 1) + isg2 - 1];
        if (fabs(ci2) < EPS6) continue;
        const double cc = ci1 * ci2;
        for (int lf1 = 0; lf1 <= in->lmax; lf1++) {
          if (in->lplus1 && lf1 < li + 1) continue;
          if (in->lmoins1 && lf1 > li) continue;
          if (lf1 > li + irang || lf1 < li - irang || lf1 % 2 != (li + irang) % 2) continue;
          for (int mf1 = -lf1; mf1 <= lf1; mf1++) {
            const int lm1 = lf1 * lf1 + lf1 + 1 + mf1;
            if (lm1 > nlm) continue;
            for (int ispf1 = 1; ispf1 <= 2; ispf1++) {
              if (ispf1 != isg1) continue;
              double gr, gi;
              orc_gaunt_xrc(lf1, mf1, l_s, m_s, li, mi1, &gr, &gi);
              const double complex gs = CMPLX(gr, gi);
              if (cabs(gs) < EPS10) continue;
              for (int lf2 = 0; lf2 <= in->lmax; lf2++) {
                if (in->lplus1 && lf2 < li + 1) continue;
                if (in->lmoins1 && lf2 > li) continue;
                if (lf2 > li + jrang || lf2 < li - jrang || lf2 % 2 != (li + jrang) % 2) continue;
                for (int mf2 = -lf2; mf2 <= lf2; mf2++) {
                  const int lm2 = lf2 * lf2 + lf2 + 1 + mf2;
                  if (lm2 > nlm) continue;
                  for (int ispf2 = 1; ispf2 <= 2; ispf2++) {
                    if (ispf2 != isg2) continue;
                    if (ispf2 != ispf1) continue;
                    orc_gaunt_xrc(lf2, mf2, l_i, m_i, li, mi2, &gr, &gi);
                    const double complex gii = CMPLX(gr, gi);
                    if (cabs(gii) < EPS10) continue;
                    const double complex cg = cc * conj(gs) * gii;
                    const double complex rof1 = rof[((size_t)(iseuil - 1) * nip + irang - in->ip0) * nlm + lm1 - 1];
                    const double complex rof2 = rof[((size_t)(iseuil - 1) * nip + jrang - in->ip0) * nlm + lm2 - 1];
                    const double complex cfe = conj(rof1) * rof2 * taull[(size_t)(lm1 - 1) * nlm + lm2 - 1];
                    const double complex cfs = conj(rof2) * rof1 * taull[(size_t)(lm2 - 1) * nlm + lm1 - 1];
                    const double complex tau_rad = 0.5 * (cfe - conj(cfs));
                    Ten[initr - 1] = Ten[initr - 1] + cg * tau_rad;
                  }
                }
              }
            }
          }
        }
      }
    }
  }
  /* times i, so that the real part is the absorption (tensor.f90:1326-1327) */
  for (int k = 0; k < ninitr; k++) Ten[k] = I * Ten[k];
}

/* The transformation coefficients Cartesian -> spherical of the dipole and
 * quadrupole operators, tenseur_car tensor.f90:240-275 (lm 0-based). */
static void clm_table(double clm[3][3][3][16]) {
  memset(clm, 0, sizeof(double) * 3 * 3 * 3 * 16);
  const double c1 = sqrt(4 * PI / 3);
  for (int j = 0; j < 3; j++) {
    clm[1][0][j][3] = c1;
    clm[1][1][j][1] = c1;
    clm[1][2][j][2] = c1;
  }
  const double c0 = sqrt(4 * PI) / 3, c = sqrt(4 * PI / 15), c3 = c / sqrt(3.0);
  clm[2][0][0][0] = c0; clm[2][0][0][6] = -c3; clm[2][0][0][8] = c;
  clm[2][0][1][4] = c;
  clm[2][0][2][7] = c;
  clm[2][1][1][0] = c0; clm[2][1][1][6] = -c3; clm[2][1][1][8] = -c;
  clm[2][1][2][5] = c;
  clm[2][2][2][0] = c0; clm[2][2][2][6] = 2 * c3;
  for (int i = 0; i < 3; i++)
    for (int j = i + 1; j < 3; j++)
      for (int lm = 0; lm < 16; lm++) clm[2][j][i][lm] = clm[2][i][j][lm];
}

#define I3(a, b) ((a) * 3 + (b))
#define I4(a, b, c) (((a) * 3 + (b)) * 3 + (c))
#define I5(a, b, c, d) ((((a) * 3 + (b)) * 3 + (c)) * 3 + (d))

/* tenseur_car, tensor.f90:15-946, electric dipole and quadrupole tensors
 * (E1E1, E1E2, E2E2; n_rel = 1, not Green_int): for every independent
 * Cartesian element (ldip, lqua, the msym tables) the sum of clm x clm x
 * tens_ab over the spherical components, spread to the elements equal by
 * symmetry, then turned by rot_int x rot_atom_abs^T.  Outputs complex
 * interleaved, C order: secdd[initr][k_s][k_i], secdq[initr][k_s][k_i][j_i],
 * secqq[initr][k_s][j_s][k_i][j_i]. */
int orc_tenseur(const orc_tens_in *in, const double *rof_re, const double *rof_im, const double *taull_re,
                const double *taull_im, double *secdd_out, double *secdq_out, double *secqq_out) {
  const int nlm = in->nlm_probe, nip = in->ip_max - in->ip0 + 1;
  const int ninitr = in->core_resolved ? in->ninit : in->nbseuil;
  if (in->lmax > 3 || nlm > 16 || ninitr > 16 || in->ip0 > 1 || in->ip_max < ((in->e1e2 || in->e2e2) ? 2 : 1))
    return -1;
  double complex *rof = malloc(sizeof(double complex) * (size_t)in->nbseuil * nip * nlm);
  double complex *taull = malloc(sizeof(double complex) * (size_t)nlm * nlm);
  for (size_t k = 0; k < (size_t)in->nbseuil * nip * nlm; k++) rof[k] = CMPLX(rof_re[k], rof_im[k]);
  for (size_t k = 0; k < (size_t)nlm * nlm; k++) taull[k] = CMPLX(taull_re[k], taull_im[k]);
  double clm[3][3][3][16];
  clm_table(clm);
  double complex *dd = calloc((size_t)ninitr * 9, sizeof(double complex));
  double complex *dq = calloc((size_t)ninitr * 27, sizeof(double complex));
  double complex *qq = calloc((size_t)ninitr * 81, sizeof(double complex));
  double complex Tens_lm[9][9][16], Ten[16], Tens[16];
  const int nrang = (in->e1e2 || in->e2e2) ? 2 : 1;
  for (int irang = 1; irang <= nrang; irang++)
    for (int jrang = irang; jrang <= nrang; jrang++) {
      memset(Tens_lm, 0, sizeof(Tens_lm));
      if (!in->e1e1 && irang == 1 && jrang == 1) continue;
      if (!in->e1e2 && irang == 1 && jrang == 2) continue;
      if (!in->e2e2 && irang == 2 && jrang == 2) continue;
      for (int ks = 0; ks < 3; ks++) {
        if (irang == 1 && in->ldip[ks] != 1) continue;
        const int njs = irang > 1 ? 3 : 1;
        for (int js = 0; js < njs; js++) {
          if (irang == 2 && in->lqua[I3(ks, js)] != 1) continue;
          for (int ki = 0; ki < 3; ki++) {
            if (jrang == 1 && in->ldip[ki] != 1) continue;
            const int nji = jrang > 1 ? 3 : 1;
            for (int ji = 0; ji < nji; ji++) {
              if (jrang == 2 && in->lqua[I3(ki, ji)] != 1) continue;
              double sa = 0;
              if (irang == 1 && jrang == 1) {
                if (ki < ks) continue;
                if (in->msymdd[I3(ks, ki)] == 0 && in->msymddi[I3(ks, ki)] == 0) continue;
                for (int k = 0; k < ninitr; k++) sa += cabs(dd[k * 9 + I3(ks, ki)]);
              } else if (irang == 1 && jrang == 2) {
                if (in->msymdq[I4(ks, ki, ji)] == 0 && in->msymdqi[I4(ks, ki, ji)] == 0) continue;
                for (int k = 0; k < ninitr; k++) sa += cabs(dq[k * 27 + I4(ks, ki, ji)]);
              } else {
                if (in->msymqq[I5(ks, js, ki, ji)] == 0 && in->msymqqi[I5(ks, js, ki, ji)] == 0) continue;
                for (int k = 0; k < ninitr; k++) sa += cabs(qq[k * 81 + I5(ks, js, ki, ji)]);
              }
              if (sa > 1e-15) continue;
              for (int k = 0; k < ninitr; k++) Tens[k] = 0;
              int lms = 0;
              for (int l_s = 0; l_s <= irang; l_s++)
                for (int m_s = -l_s; m_s <= l_s; m_s++, lms++) {
                  const double clm_s = clm[irang][ks][js][lms];
                  if (fabs(clm_s) < EPS10) continue;
                  int lmi = 0;
                  for (int l_i = 0; l_i <= jrang; l_i++)
                    for (int m_i = -l_i; m_i <= l_i; m_i++, lmi++) {
                      const double clm_i = clm[jrang][ki][ji][lmi];
                      if (fabs(clm_i) < EPS10) continue;
                      double st = 0;
                      for (int k = 0; k < ninitr; k++) st += cabs(Tens_lm[lms][lmi][k]);
                      if (st < 1e-15) {
                        tens_ab_r(in, ninitr, irang, jrang, l_s, m_s, l_i, m_i, rof, taull, Ten);
                        if (in->numat == 1)
                          for (int k = 0; k < ninitr; k++) Ten[k] = Ten[k] / 2;
                        for (int k = 0; k < ninitr; k++) Tens_lm[lms][lmi][k] = Ten[k];
                      }
                      for (int k = 0; k < ninitr; k++) Tens[k] = Tens[k] + clm_s * clm_i * Tens_lm[lms][lmi][k];
                    }
                }
              /* spread to the elements equal by symmetry (tensor.f90:529-700) */
              if (irang == 1 && jrang == 1) {
                const int a0 = in->msymdd[I3(ks, ki)], b0 = in->msymddi[I3(ks, ki)];
                for (int kks = 0; kks < 3; kks++)
                  for (int kki = kks; kki < 3; kki++) {
                    const int a = in->msymdd[I3(kks, kki)], b = in->msymddi[I3(kks, kki)];
                    if (abs(a) != abs(a0) || abs(b) != abs(b0)) continue;
                    const int is = a != 0 ? a / a0 : 0, isi = b != 0 ? b / b0 : 0;
                    for (int k = 0; k < ninitr; k++) {
                      const double complex te = CMPLX(is * creal(Tens[k]), isi * cimag(Tens[k]));
                      dd[k * 9 + I3(kks, kki)] = te;
                      dd[k * 9 + I3(kki, kks)] = conj(te);
                    }
                  }
              } else if (irang == 1 && jrang == 2) {
                const int a0 = in->msymdq[I4(ks, ki, ji)], b0 = in->msymdqi[I4(ks, ki, ji)];
                for (int kks = 0; kks < 3; kks++)
                  for (int kki = 0; kki < 3; kki++)
                    for (int jji = kki; jji < 3; jji++) {
                      const int a = in->msymdq[I4(kks, kki, jji)], b = in->msymdqi[I4(kks, kki, jji)];
                      if (abs(a) != abs(a0) || abs(b) != abs(b0)) continue;
                      const int is = a != 0 ? a / a0 : 0, isi = b != 0 ? b / b0 : 0;
                      for (int k = 0; k < ninitr; k++) {
                        const double complex te = CMPLX(is * creal(Tens[k]), isi * cimag(Tens[k]));
                        dq[k * 27 + I4(kks, kki, jji)] = te;
                        dq[k * 27 + I4(kks, jji, kki)] = te;
                      }
                    }
              } else {
                const int a0 = in->msymqq[I5(ks, js, ki, ji)], b0 = in->msymqqi[I5(ks, js, ki, ji)];
                for (int kks = 0; kks < 3; kks++)
                  for (int jjs = 0; jjs < 3; jjs++)
                    for (int kki = 0; kki < 3; kki++)
                      for (int jji = 0; jji < 3; jji++) {
                        double s2 = 0;
                        for (int k = 0; k < ninitr; k++) s2 += cabs(qq[k * 81 + I5(kks, jjs, kki, jji)]);
                        if (s2 > 1e-15) continue;
                        const int a = in->msymqq[I5(kks, jjs, kki, jji)], b = in->msymqqi[I5(kks, jjs, kki, jji)];
                        if (abs(a) != abs(a0) || abs(b) != abs(b0)) continue;
                        const int is = a != 0 ? a / a0 : 0, isi = b != 0 ? b / b0 : 0;
                        for (int k = 0; k < ninitr; k++) {
                          double complex te = CMPLX(is * creal(Tens[k]), isi * cimag(Tens[k]));
                          double complex *q = qq + k * 81;
                          q[I5(kks, jjs, kki, jji)] = te;
                          q[I5(jjs, kks, kki, jji)] = te;
                          q[I5(kks, jjs, jji, kki)] = te;
                          q[I5(jjs, kks, jji, kki)] = te;
                          te = conj(te);
                          q[I5(kki, jji, kks, jjs)] = te;
                          q[I5(jji, kki, kks, jjs)] = te;
                          q[I5(kki, jji, jjs, kks)] = te;
                          q[I5(jji, kki, jjs, kks)] = te;
                        }
                      }
              }
            }
          }
        }
      }
    }
  /* rotation to the R1 basis (tensor.f90:781-806; rot_tensor_2/3/4, :2112-2210) */
  double rot[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double t = 0;
      for (int k = 0; k < 3; k++) t += in->rot_int[I3(i, k)] * in->rot_atom_abs[I3(j, k)];
      rot[i][j] = t;
    }
  for (int k = 0; k < ninitr; k++) {
    if (in->e1e1) {
      const double complex *m2 = dd + k * 9;
      double complex out[9] = {0};
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          for (int kk = 0; kk < 3; kk++) {
            double complex cmat = 0;
            for (int l = 0; l < 3; l++) cmat += rot[j][l] * m2[I3(kk, l)];
            out[I3(i, j)] = out[I3(i, j)] + rot[i][kk] * cmat;
          }
      memcpy(dd + k * 9, out, sizeof(out));
    }
    if (in->e1e2) {
      const double complex *m3 = dq + k * 27;
      double complex out[27] = {0};
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          for (int kk = 0; kk < 3; kk++)
            for (int l = 0; l < 3; l++) {
              double complex cmas = 0;
              for (int m = 0; m < 3; m++) {
                double complex cmat = 0;
                for (int n = 0; n < 3; n++) cmat += rot[kk][n] * m3[I4(l, m, n)];
                cmas = cmas + rot[j][m] * cmat;
              }
              out[I4(i, j, kk)] = out[I4(i, j, kk)] + rot[i][l] * cmas;
            }
      memcpy(dq + k * 27, out, sizeof(out));
    }
    if (in->e2e2) {
      const double complex *m4 = qq + k * 81;
      double complex out[81] = {0};
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          for (int kk = 0; kk < 3; kk++)
            for (int l = 0; l < 3; l++)
              for (int m = 0; m < 3; m++) {
                double complex cmar = 0;
                for (int n = 0; n < 3; n++) {
                  double complex cmas = 0;
                  for (int n1 = 0; n1 < 3; n1++) {
                    double complex cmat = 0;
                    for (int p = 0; p < 3; p++) cmat += rot[l][p] * m4[I5(m, n, n1, p)];
                    cmas = cmas + rot[kk][n1] * cmat;
                  }
                  cmar = cmar + rot[j][n] * cmas;
                }
                out[I5(i, j, kk, l)] = out[I5(i, j, kk, l)] + rot[i][m] * cmar;
              }
      memcpy(qq + k * 81, out, sizeof(out));
    }
  }
  for (int k = 0; k < ninitr * 9; k++) { secdd_out[2 * k] = creal(dd[k]); secdd_out[2 * k + 1] = cimag(dd[k]); }
  for (int k = 0; k < ninitr * 27; k++) { secdq_out[2 * k] = creal(dq[k]); secdq_out[2 * k + 1] = cimag(dq[k]); }
  for (int k = 0; k < ninitr * 81; k++) { secqq_out[2 * k] = creal(qq[k]); secqq_out[2 * k + 1] = cimag(qq[k]); }
  free(rof); free(taull); free(dd); free(dq); free(qq);
  return ninitr;
}
#undef I3
#undef I4
#undef I5
