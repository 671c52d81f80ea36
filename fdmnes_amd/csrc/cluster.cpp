// Synthetic FDM cluster builder (host).  See include/fdmgpu_cluster.h for the
// reference routines each step stands in for.
#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/fdmgpu_cluster.h"
#include "ylm.h"

namespace {

typedef std::complex<double> cd;
constexpr double kPi = 3.1415926535897932384626433832;
constexpr double kAlfaSf = 0.0072973525664;  // fine-structure constant (main.f90:58)
constexpr double kEpsPos = 1e-4;  // tolerance on point positions (main.f90:83)

struct GridPoint {
  int ix, iy, iz;
};

// Stencil directions in the reference's neighbour order (laplac,
// general.f90:9760-9772): distance io = 1, 2; axis x, y, z; side -, +.
const int kDir[12][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1},
                         {-2, 0, 0}, {2, 0, 0}, {0, -2, 0}, {0, 2, 0}, {0, 0, -2}, {0, 0, 2}};

using fdmgpu::real_ylm;

// Spherical Bessel and Neumann functions times fnorm, by upward recurrence
// (the published algorithm of sphere.f90:7226-7248): bes_l = fnorm j_l(z),
// neu_l = fnorm y_l(z).
template <class T>
void bessel_neumann(T fnorm, T z, int lmax, T *bes, T *neu) {
  neu[0] = -fnorm * std::cos(z) / z;
  bes[0] = fnorm * std::sin(z) / z;
  if (lmax < 1) return;
  neu[1] = neu[0] / z - bes[0];
  bes[1] = bes[0] / z + neu[0];
  for (int l = 2; l <= lmax; l++) {
    neu[l] = T(2 * l - 1) * neu[l - 1] / z - neu[l - 2];
    bes[l] = T(2 * l - 1) * bes[l - 1] / z - bes[l - 2];
  }
}

}  // namespace

struct fdmgpu_cluster {
  double rsort = 0, adimp = 0;
  int lmaxso = 0, nlmso = 0;
  // -1: the whole cluster (point group P1).  Otherwise one representation of
  // the group of the mirrors in `mirrors` (bit a = the mirror normal to axis
  // a; 4 = m, 3 = mm2, 7 = mmm): bit a of irrep set = odd under that mirror
  // (a subset of mirrors); the grid is then the wedge with coordinate a >= 0
  // on every mirror axis
  int irrep = -1;
  int mirrors = 0;
  // The point group as axis sign flips (bit a: coordinate a negated): every
  // group built here is a subgroup of mmm.  group: the reference's number
  // (general.f90:5207-5220): 1, -1 = 2, m = 3, 2 = 4, 2/m = 5, mm2 = 6,
  // 222 = 7, mmm = 8.  ops: its operations in the reference's order of the
  // symmetry operations (opsym: identity 1, two-fold axes x, y, z 22-24,
  // inversion 25, mirrors normal to x, y, z 40-42), as flip masks.  For the
  // groups with rotations or inversion the representation bits irrep are a
  // character vector: the character of the operation with flips g is
  // (-1)^popcount(g & irrep); irrep is the smallest mask giving the
  // representation's characters on the group.
  int group = 1;
  std::vector<int> ops{0};
  int npoint = 0, npso = 0, nligne = 0;
  std::vector<GridPoint> pt;   // npso, radial order
  std::vector<double> rad;     // npso, |index| * adimp
  std::vector<int> nbr;        // npso * 12, 0-based point or -1
  std::vector<unsigned char> flip;  // npso * 12: mirrors taking the neighbour into the wedge (bit per axis)
  std::vector<double> rvol;    // npso: points of the whole grid a wedge point stands for, over the group order (1 for P1)
  std::vector<double> wso;     // nsortf: outer-sphere quadrature weights 4 pi rvol / sum(rvol)
  std::vector<int> lmidx;      // nlmso: harmonic index l*l + l + m of each outer unknown
  std::vector<int> bnd;        // boundary points: inner (nsortf) then outer
  int nsortf = 0;
  std::vector<int> slot;       // per point: index in bnd of an outer point, -1 otherwise
  std::vector<double> ylm;     // bnd.size() * nlmso
  std::vector<int> lb1, lb2;   // nligne, 1-based
  std::vector<int64_t> rp;     // compressed-row structure, nligne + 1
  std::vector<int> cols;       // 1-based
  double clapl[13] = {0};
  double cgrad[13] = {0};  // gradient stencil (laplac), 1..12 in kDir order
  // Muffin-tin atoms (point group P1 only): positions (bohr, absorber at
  // the origin, sorted by distance), sphere radii, expansions
  int natome = 0, npr = 0;
  std::vector<double> apos, armt;
  std::vector<int> almax, nlmsa, ianew;  // ianew: rows ianew .. ianew + nlmsa - 1 (0-based) are the atom's
  std::vector<int> numia;                // npoint: 0, or 1 + the atom the point lies in
  std::vector<int> row;                  // npoint: 0-based row of an interstitial point, -1 inside an atom
  std::vector<std::vector<int>> abord;   // per atom: nbordf outside border points, then the inside ones
  std::vector<int> nbordf;
  std::vector<std::vector<double>> apoids;  // per atom, nbordf weights 4 pi rvol / sum(rvol)
  std::vector<std::vector<double>> aylm;    // per atom, nbord x nlmsa real harmonics about the atom
  std::vector<int> rkind, ridx;          // per row: 0 grid (point), 1 atom (atom, Lm packed), 2 outer (lm)
  int nso = 0;                           // first outer row (0-based)
  // the irregular cluster solution (mat.f90:140-147, Test_reg = 0): two
  // blocks per atom (regular then irregular atomic solutions, nri = 2), no
  // outer-sphere rows (nso = nligne)
  bool irregular = false;
  int nri = 1;
  // complex harmonics in the outer sphere (the reference's Ylm_comp): the
  // outer unknowns and incoming waves are Yc(L, m) (mat.f90:1733), with
  // their values at the outer points in ycmp (bnd.size() x nlmso)
  bool ylm_comp = false;
  std::vector<cd> ycmp;
  // the groups 4, -4 and 4/m (numbers 9-11), whose four-fold operations are
  // no axis sign flips: the operations as matrices (row-major, p -> M p) in
  // the reference's order, per neighbour the one that maps it into the
  // calculation area, and the representation's (complex) characters
  std::vector<std::array<int, 9>> gmat;
  std::vector<unsigned char> gop;
  std::vector<cd> gchar;
};

namespace {

// Parities of the real harmonic l*l + l + m under the three mirrors, as the
// representation bits: cos(m phi) P_l^m (m >= 0) is (-1)^m in x, even in y,
// (-1)^(l+m) in z; sin(|m| phi) P_l^|m| is (-1)^(|m|+1) in x, odd in y,
// (-1)^(l+|m|) in z.
int harmonic_irrep(int l, int m) {
  const int am = m < 0 ? -m : m;
  const int ox = (m >= 0) ? (am & 1) : ((am + 1) & 1);
  const int oy = (m < 0) ? 1 : 0;
  const int oz = (l + am) & 1;
  return ox | (oy << 1) | (oz << 2);
}

// character of the mirrors in mask for representation g: -1 per odd axis flipped
inline double character(int g, int mask) { return (__builtin_popcount(g & mask) & 1) ? -1.0 : 1.0; }

// The calculation area of each group, posequiv's regions (general.f90:
// 8950-8978) on integer grid coordinates
bool in_domain(int group, int x, int y, int z) {
  switch (group) {
    case 2: return z > 0 || (z == 0 && (x > 0 || (x == 0 && y >= 0)));  // -1
    case 3: return z >= 0;                                               // m
    case 4: return x > 0 || (x == 0 && y >= 0);                          // 2 (axis z)
    case 5: return z >= 0 && (x > 0 || (x == 0 && y >= 0));              // 2/m
    case 6: return x >= 0 && y >= 0;                                     // mm2
    case 7: return (x > 0 && y > 0) || (x == 0 && y >= 0 && z >= 0) || (y == 0 && x >= 0 && z >= 0);  // 222
    case 8: return x >= 0 && y >= 0 && z >= 0;                           // mmm
    case 9: return (x > 0 && y >= 0) || (x == 0 && y == 0);              // 4
    case 10: return (x > 0 && y >= 0) || (x == 0 && y == 0 && z >= 0);   // -4
    case 11: return z >= 0 && ((x > 0 && y >= 0) || (x == 0 && y == 0));  // 4/m
    default: return true;
  }
}

// group number -> its operations as flip masks, in opsym order:
// 1, 22 (2x: y z), 23 (2y: x z), 24 (2z: x y), 25 (-1: x y z), 40 (m_x),
// 41 (m_y), 42 (m_z)
std::vector<int> group_ops(int group) {
  switch (group) {
    case 1: return {0};
    case 2: return {0, 7};
    case 3: return {0, 4};
    case 4: return {0, 3};
    case 5: return {0, 3, 7, 4};
    case 6: return {0, 3, 1, 2};
    case 7: return {0, 6, 5, 3};
    case 8: return {0, 6, 5, 3, 7, 1, 2, 4};
    default: return {};
  }
}

inline bool is_mirror_group(int group) { return group == 3 || group == 6 || group == 8; }
inline bool is_c4_group(int group) { return group >= 9 && group <= 11; }

// The operations of 4 (9), -4 (10) and 4/m (11) in the reference's order
// (opsym, general.f90:4587; group tables tab_data.f90:715), as opsym's
// matrices act on a position: identity 1, C4z 18: (-y, x, z), -C4z 21:
// (y, -x, z), C2z 24, inversion 25, S4z 28: (-y, x, -z), -S4z 31:
// (y, -x, -z), mirror z 42
std::vector<std::array<int, 9>> c4_ops(int group) {
  const std::array<int, 9> e{1, 0, 0, 0, 1, 0, 0, 0, 1}, c4{0, -1, 0, 1, 0, 0, 0, 0, 1},
      c4m{0, 1, 0, -1, 0, 0, 0, 0, 1}, c2{-1, 0, 0, 0, -1, 0, 0, 0, 1}, inv{-1, 0, 0, 0, -1, 0, 0, 0, -1},
      s4{0, -1, 0, 1, 0, 0, 0, 0, -1}, s4m{0, 1, 0, -1, 0, 0, 0, 0, -1}, mz{1, 0, 0, 0, 1, 0, 0, 0, -1};
  switch (group) {
    case 9: return {e, c4, c4m, c2};
    case 10: return {e, c2, s4, s4m};
    case 11: return {e, c4, c4m, c2, inv, s4, s4m, mz};
    default: return {};
  }
}
// i^e exactly
inline cd ipow(int e) {
  e = ((e % 4) + 4) % 4;
  return cd(e == 0 ? 1.0 : (e == 2 ? -1.0 : 0.0), e == 1 ? 1.0 : (e == 3 ? -1.0 : 0.0));
}
// quarter turns q of an operation's rotation about z: (cos, sin) of q pi/2 = (M00, M10)
inline int quarter_turns(const std::array<int, 9> &M) {
  return M[0] == 1 ? 0 : (M[3] == 1 ? 1 : (M[0] == -1 ? 2 : 3));
}
// Character of the operation in representation v = k + 4 [p = -1]: the
// rotation by q pi/2 has exp(-i k q pi/2), a reflection of z adds p (p = 1
// for -4, whose representations are k = 0..3)
inline cd c4_char(const std::array<int, 9> &M, int v) {
  const int k = v & 3;
  cd chi = ipow(-k * quarter_turns(M));
  if (M[8] < 0 && (v & 4)) chi = -chi;
  return chi;
}
// The complex harmonic Yc(l, m) (exp(i m phi) dependence) belongs to v when
// Y(g r) = chi(g) Y(r) for every operation g (the neighbour relation
// psi(n) = chi(g^-1) psi(g n) of the grid rows): a rotation by q pi/2
// multiplies it by i^(m q), a reflection of z by (-1)^(l+m)
bool c4_harmonic_in(const std::vector<std::array<int, 9>> &ops, int l, int m, int v) {
  for (const auto &M : ops) {
    cd lam = ipow(m * quarter_turns(M));
    if (M[8] < 0 && ((l + m) & 1)) lam = -lam;
    if (c4_char(M, v) != lam) return false;
  }
  return true;
}
inline void apply(const std::array<int, 9> &M, int x, int y, int z, int &u, int &v, int &w) {
  u = M[0] * x + M[1] * y + M[2] * z;
  v = M[3] * x + M[4] * y + M[5] * z;
  w = M[6] * x + M[7] * y + M[8] * z;
}
// Yc, mat.f90:1733-1755: the complex harmonic (l, m) from the real ones
// y1 = Y(l, m), y2 = Y(l, -m)
inline cd yc(int m, double y1, double y2) {
  const double r = 1 / 1.41421356237309504880168872420970;
  if (m == 0) return cd(y1, 0.0);
  if (m < 0) return cd(y2 * r, -y1 * r);
  const double sg = (m & 1) ? -1.0 : 1.0;
  return cd(y1 * sg * r, y2 * sg * r);
}

// Does the real harmonic with flip parities o belong to representation v of
// the group: its parity under every operation g equals the character of g
bool harmonic_in(const std::vector<int> &ops, int o, int v) {
  for (int g : ops)
    if (__builtin_popcount(g & (o ^ v)) & 1) return false;
  return true;
}

inline int lof(int lm) {  // l of the harmonic index l*l + l + m
  int l = 0;
  while ((l + 1) * (l + 1) <= lm) l++;
  return l;
}

// The atom border lists, harmonics and weights (bordure, general.f90:
// 9924-10068; Ylmpt :11533-11568), the rows with the atom blocks in the
// radial order (newind, mat.f90:1790-1884) and the compressed-row structure.
int build_atoms(fdmgpu_cluster *c) {
  const int np = c->npoint, na = c->natome, nl = c->nlmso;
  const GridPoint *p = c->pt.data();
  c->abord.assign(na, {});
  c->nbordf.assign(na, 0);
  std::vector<int> ibp(na, -1);
  for (int i = 0; i < np; i++) {  // outside border points: a first-shell neighbour in the atom
    if (c->numia[i]) continue;
    for (int v = 0; v < 6; v++) {
      const int j = c->nbr[(size_t)i * 12 + v];
      if (j < 0 || j >= np || !c->numia[j]) continue;
      const int a = c->numia[j] - 1;
      if (ibp[a] != i) c->abord[a].push_back(i);
      ibp[a] = i;
    }
  }
  for (int a = 0; a < na; a++) c->nbordf[a] = (int)c->abord[a].size();
  for (int i = 0; i < np; i++) {  // inside border points: a neighbour in the interstitial grid
    if (!c->numia[i]) continue;
    for (int v = 0; v < 12; v++) {
      const int j = c->nbr[(size_t)i * 12 + v];
      if (j < 0) return -1;
      if (j >= np || c->numia[j]) continue;
      c->abord[c->numia[i] - 1].push_back(i);
      break;
    }
  }
  c->apoids.assign(na, {});
  c->aylm.assign(na, {});
  c->nlmsa.resize(na);
  for (int a = 0; a < na; a++) {
    if (c->nbordf[a] == 0) return -1;  // an atom the grid does not see
    double pd = 0;
    for (int b = 0; b < c->nbordf[a]; b++) pd += c->rvol[c->abord[a][b]];
    for (int b = 0; b < c->nbordf[a]; b++) c->apoids[a].push_back(4 * kPi * c->rvol[c->abord[a][b]] / pd);
    const int lm = (c->almax[a] + 1) * (c->almax[a] + 1);
    c->nlmsa[a] = lm;
    c->aylm[a].resize(c->abord[a].size() * lm);
    for (size_t b = 0; b < c->abord[a].size(); b++) {
      const GridPoint &q = p[c->abord[a][b]];
      real_ylm(c->almax[a], q.ix * c->adimp - c->apos[3 * a], q.iy * c->adimp - c->apos[3 * a + 1],
               q.iz * c->adimp - c->apos[3 * a + 2], &c->aylm[a][b * lm]);
    }
  }
  // rows: interstitial points in order, each atom's block before the first
  // interstitial point at least as far from the absorber as the atom
  std::vector<int> nw(np + 1, 0);  // 1-based rows of the reference's new(i)
  int ii = 0, npr = 0;
  for (int i = 0; i < np; i++)
    if (!c->numia[i]) nw[i] = ++ii, npr++;
  c->npr = npr;
  c->ianew.assign(na, 0);
  int npf = 0, npf1 = -1;
  for (int a = 0; a < na; a++) {
    const double dist = std::sqrt(c->apos[3 * a] * c->apos[3 * a] + c->apos[3 * a + 1] * c->apos[3 * a + 1] +
                                  c->apos[3 * a + 2] * c->apos[3 * a + 2]);
    int i = npf;
    while (i < np && !(c->numia[i] == 0 && c->rad[i] >= dist)) i++;
    if (i >= np) return -1;  // no grid row beyond the atom
    npf = i;
    c->ianew[a] = (npf1 == npf) ? c->ianew[a - 1] + c->nri * c->nlmsa[a - 1] : nw[npf] - 1;
    npf1 = npf;
    for (int k = npf; k < np; k++)
      if (!c->numia[k]) nw[k] += c->nri * c->nlmsa[a];
  }
  int nat = 0;
  for (int a = 0; a < na; a++) nat += c->nri * c->nlmsa[a];
  const int nout = c->irregular ? 0 : nl;  // no scattered wave in the irregular solution
  const int nlig = npr + nat + nout;
  c->nligne = nlig;
  c->nso = nlig - nout;
  c->row.assign(np, -1);
  for (int i = 0; i < np; i++)
    if (!c->numia[i]) c->row[i] = nw[i] - 1;
  c->rkind.assign(nlig, 0);
  c->ridx.assign(nlig, 0);
  for (int i = 0; i < np; i++)
    if (c->row[i] >= 0) c->ridx[c->row[i]] = i;
  for (int a = 0; a < na; a++)
    for (int lm = 0; lm < c->nri * c->nlmsa[a]; lm++) {  // lm >= nlmsa: the irregular block's rows
      c->rkind[c->ianew[a] + lm] = 1;
      c->ridx[c->ianew[a] + lm] = (a << 16) | lm;
    }
  for (int lm = 0; lm < nout; lm++) {
    c->rkind[c->nso + lm] = 2;
    c->ridx[c->nso + lm] = lm;
  }
  // profile (newind, mat.f90:1812-1884), 1-based
  c->lb1.resize(nlig);
  c->lb2.resize(nlig);
  for (int r = 1; r <= nlig; r++) c->lb1[r - 1] = c->lb2[r - 1] = r;
  for (int i = 0; i < np; i++) {
    if (c->row[i] < 0) continue;
    int &l1 = c->lb1[c->row[i]], &l2 = c->lb2[c->row[i]];
    for (int v = 0; v < 12; v++) {
      const int j = c->nbr[(size_t)i * 12 + v];
      if (j < 0) return -1;
      if (j >= np) {
        l2 = nlig;
      } else if (c->numia[j]) {
        const int a = c->numia[j] - 1;
        l1 = std::min(l1, c->ianew[a] + 1);
        l2 = std::max(l2, c->ianew[a] + c->nri * c->nlmsa[a]);
      } else {
        l1 = std::min(l1, c->row[j] + 1);
        l2 = std::max(l2, c->row[j] + 1);
      }
    }
  }
  for (int a = 0; a < na; a++) {
    int l1 = c->ianew[a] + 1, l2 = c->ianew[a] + c->nri * c->nlmsa[a];
    for (int b = 0; b < c->nbordf[a]; b++) {
      l1 = std::min(l1, c->row[c->abord[a][b]] + 1);
      l2 = std::max(l2, c->row[c->abord[a][b]] + 1);
    }
    for (int lm = 0; lm < c->nri * c->nlmsa[a]; lm++) {
      c->lb1[c->ianew[a] + lm] = l1;
      c->lb2[c->ianew[a] + lm] = l2;
    }
  }
  if (!c->irregular) {
    int lo = c->nso + 1;
    for (int b = 0; b < c->nsortf; b++) lo = std::min(lo, c->row[c->bnd[b]] + 1);
    for (int r = c->nso + 1; r <= nlig; r++) {
      c->lb1[r - 1] = lo;
      c->lb2[r - 1] = nlig;
    }
  }
  for (int r = nlig - 1; r >= 1; r--) c->lb1[r - 1] = std::min(c->lb1[r - 1], c->lb1[r]);
  // compressed rows, columns ascending: grid rows (diagonal, interstitial
  // neighbours, the blocks of atoms a neighbour lies in, outer columns),
  // atom rows (the outside border points, the atom's block), outer rows
  c->rp.assign(nlig + 1, 0);
  c->cols.clear();
  std::vector<int> tmp;
  for (int r = 0; r < nlig; r++) {
    tmp.clear();
    if (c->rkind[r] == 0) {
      const int i = c->ridx[r];
      tmp.push_back(r + 1);
      bool outside = false;
      for (int v = 0; v < 12; v++) {
        const int j = c->nbr[(size_t)i * 12 + v];
        if (j >= np) {
          outside = true;
        } else if (c->numia[j]) {
          const int a = c->numia[j] - 1;
          for (int lm = 0; lm < c->nri * c->nlmsa[a]; lm++) tmp.push_back(c->ianew[a] + lm + 1);
        } else {
          tmp.push_back(c->row[j] + 1);
        }
      }
      if (outside)
        for (int l = 0; l < nout; l++) tmp.push_back(c->nso + l + 1);
    } else if (c->rkind[r] == 1) {
      const int a = c->ridx[r] >> 16;
      // every outside border point, as mat_solve's symbolic pass marks them
      // (mat_solve_gaussian.f90:123-137); in the irregular solution each row
      // block fills only its half of them (mat.f90:1553-1558)
      for (int b = 0; b < c->nbordf[a]; b++) tmp.push_back(c->row[c->abord[a][b]] + 1);
      for (int k = 0; k < c->nri * c->nlmsa[a]; k++) tmp.push_back(c->ianew[a] + k + 1);
    } else {
      for (int b = 0; b < c->nsortf; b++) tmp.push_back(c->row[c->bnd[b]] + 1);
      for (int k = 0; k < nl; k++) tmp.push_back(c->nso + k + 1);
    }
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    c->cols.insert(c->cols.end(), tmp.begin(), tmp.end());
    c->rp[r + 1] = (int64_t)c->cols.size();
  }
  return 0;
}

int build(fdmgpu_cluster *c) {
  const bool sym = c->irrep >= 0;
  const int mx = sym ? c->mirrors : 0;
  // Grid points inside rsort/adimp + sqrt(2)*iord/2 (reseau, general.f90:9476)
  // visited in (ix, iy, iz) lexicographic order, then ordered by radius with
  // the reference's exchange sort (:9528-9543): a later point is swapped in
  // only if it is closer by more than epspos, so equal radii keep the
  // arrival order this exchange pattern produces.
  const double rmax = c->rsort / c->adimp + 1.41421356237309504880168872420970 * 4 / 2 + kEpsPos;
  const int nx = (int)std::ceil(rmax) + 1;
  for (int ix = -nx; ix <= nx; ix++)
    for (int iy = -nx; iy <= nx; iy++)
      for (int iz = -nx; iz <= nx; iz++) {
        const double d = std::sqrt((double)ix * ix + (double)iy * iy + (double)iz * iz);
        if (d > rmax) continue;
        // with symmetry only the calculation area (posequiv's region of the
        // group, general.f90:8950-8978) is kept (reseau, general.f90:9512)
        if (sym && !in_domain(c->group, ix, iy, iz)) continue;
        c->pt.push_back({ix, iy, iz});
        c->rad.push_back(c->adimp * d);
      }
  const int n = (int)c->pt.size();
  double *r = c->rad.data();
  GridPoint *p = c->pt.data();
  for (int a = 0; a < n; a++) {
    double ra = r[a];
    for (int b = a + 1; b < n; b++) {
      if (ra < r[b] + kEpsPos) continue;
      std::swap(r[a], r[b]);
      std::swap(p[a], p[b]);
      ra = r[a];
    }
  }
  c->npso = n;
  int np = 0;
  while (np < n && !(r[np] > c->rsort + kEpsPos)) np++;
  c->npoint = np;
  // Point weights (reseau, general.f90:9582-9651): the number of grid points
  // a wedge point stands for, over the largest such number
  c->rvol.assign(n, 1.0);
  if (sym && is_c4_group(c->group)) {  // orbits under the matrices
    double rvmax = 1.0;
    for (int i = 0; i < n; i++) {
      std::vector<std::array<int, 3>> img;
      for (const auto &M : c->gmat) {
        std::array<int, 3> u;
        apply(M, p[i].ix, p[i].iy, p[i].iz, u[0], u[1], u[2]);
        if (std::find(img.begin(), img.end(), u) == img.end()) img.push_back(u);
      }
      c->rvol[i] = (double)img.size();
      rvmax = std::max(rvmax, c->rvol[i]);
    }
    for (int i = 0; i < n; i++) c->rvol[i] /= rvmax;
  } else if (sym) {
    // the size of the point's orbit under the group (every image of an
    // area point lies in the grid's symmetric sphere)
    double rvmax = 1.0;
    for (int i = 0; i < n; i++) {
      // two operations give the same image when they differ only in the
      // flips of zero coordinates
      const int zero = (p[i].ix == 0) | ((p[i].iy == 0) << 1) | ((p[i].iz == 0) << 2);
      int img[8], ni = 0;
      for (int g : c->ops) {
        const int k = g & ~zero;
        bool seen = false;
        for (int t = 0; t < ni; t++) seen |= img[t] == k;
        if (!seen) img[ni++] = k;
      }
      c->rvol[i] = (double)ni;
      rvmax = std::max(rvmax, c->rvol[i]);
    }
    for (int i = 0; i < n; i++) c->rvol[i] /= rvmax;
  }
  // Neighbour table (laplac, general.f90:9839-9871); with symmetry a
  // neighbour outside the wedge is its mirror image, and flip records the
  // mirrors (isvois)
  const int w = 2 * nx + 1;
  std::vector<int> where((size_t)w * w * w, -1);
  auto key = [&](int x, int y, int z) { return ((size_t)(x + nx) * w + (y + nx)) * w + (z + nx); };
  for (int i = 0; i < n; i++) where[key(p[i].ix, p[i].iy, p[i].iz)] = i;
  c->nbr.assign((size_t)n * 12, -1);
  c->flip.assign((size_t)n * 12, 0);
  if (is_c4_group(c->group)) c->gop.assign((size_t)n * 12, 0);
  for (int i = 0; i < n; i++)
    for (int v = 0; v < 12; v++) {
      int x = p[i].ix + kDir[v][0], y = p[i].iy + kDir[v][1], z = p[i].iz + kDir[v][2];
      if (!(std::abs(x) <= nx && std::abs(y) <= nx && std::abs(z) <= nx)) continue;
      if (sym && is_c4_group(c->group)) {
        // the first operation g that maps it into the area; posequiv
        // records g's inverse, whose character the coupling then carries
        // (general.f90:9111, isym_inv :9124)
        for (size_t g = 0; g < c->gmat.size(); g++) {
          int u, v2, w;
          apply(c->gmat[g], x, y, z, u, v2, w);
          if (in_domain(c->group, u, v2, w)) {
            size_t gi = 0;  // the inverse: the transposed matrix
            for (; gi < c->gmat.size(); gi++) {
              bool tr = true;
              for (int r = 0; r < 3; r++)
                for (int q = 0; q < 3; q++) tr &= c->gmat[gi][3 * r + q] == c->gmat[g][3 * q + r];
              if (tr) break;
            }
            c->gop[(size_t)i * 12 + v] = (unsigned char)gi;
            x = u;
            y = v2;
            z = w;
            break;
          }
        }
      } else if (sym) {
        int f = 0;
        if (is_mirror_group(c->group)) {  // the mirrors of the negative coordinates
          f = ((x < 0) | ((y < 0) << 1) | ((z < 0) << 2)) & mx;
        } else {  // the first operation, in opsym order, that takes it into the area (posequiv)
          for (int g : c->ops)
            if (in_domain(c->group, (g & 1) ? -x : x, (g & 2) ? -y : y, (g & 4) ? -z : z)) {
              f = g;
              break;
            }
        }
        c->flip[(size_t)i * 12 + v] = (unsigned char)f;
        if (f & 1) x = -x;
        if (f & 2) y = -y;
        if (f & 4) z = -z;
      }
      c->nbr[(size_t)i * 12 + v] = where[key(x, y, z)];
    }
  const double c1 = (4.0 / 3.0) / (c->adimp * c->adimp), c2 = (-1.0 / 16) * c1;  // general.f90:9754-9776
  double s = 0;
  for (int v = 1; v <= 12; v++) {
    c->clapl[v] = (v <= 6) ? c1 : c2;
    s += c->clapl[v];
  }
  c->clapl[0] = -s;
  // gradient coefficients (general.f90:9791-9822): is * coeford * cd with
  // is = -1, +1 along each axis, coeford 1 and -1/8 for the two shells
  const double cdg = (4.0 / 3.0) / (2 * c->adimp);
  for (int v = 1; v <= 12; v++) c->cgrad[v] = ((v & 1) ? -1.0 : 1.0) * (v <= 6 ? 1.0 : -1.0 / 8) * cdg;
  // Points inside an atom's sphere (reseau, general.f90:9563-9577; the
  // distance to the atom, the identity being the only operation)
  c->numia.assign(np, 0);
  for (int a = 0; a < c->natome; a++)
    for (int i = 0; i < np; i++) {
      const double dx = p[i].ix * c->adimp - c->apos[3 * a], dy = p[i].iy * c->adimp - c->apos[3 * a + 1],
                   dz = p[i].iz * c->adimp - c->apos[3 * a + 2];
      if (std::sqrt(dx * dx + dy * dy + dz * dz) < c->armt[a] + kEpsPos) c->numia[i] = a + 1;
    }
  // Outer-sphere point lists (bordure, general.f90:9918-10076): interstitial
  // points with a first-shell neighbour outside, then outside points touching
  // an interstitial point; quadrature weight 4 pi / nsortf each (unit point
  // volumes).
  for (int i = 0; i < np; i++) {
    if (c->numia[i]) continue;
    for (int v = 0; v < 6; v++) {
      const int j = c->nbr[(size_t)i * 12 + v];
      if (j >= np) {
        c->bnd.push_back(i);
        break;
      }
    }
  }
  c->nsortf = (int)c->bnd.size();
  double pdsort = 0;  // bordure, general.f90:10006, :10073-10075
  for (int b = 0; b < c->nsortf; b++) pdsort += c->rvol[c->bnd[b]];
  c->wso.resize(c->nsortf);
  for (int b = 0; b < c->nsortf; b++) c->wso[b] = 4 * kPi * c->rvol[c->bnd[b]] / pdsort;
  c->slot.assign(n, -1);
  for (int i = np; i < n; i++)
    for (int v = 0; v < 12; v++) {
      const int j = c->nbr[(size_t)i * 12 + v];
      if (j >= 0 && j < np && !c->numia[j]) {
        c->slot[i] = (int)c->bnd.size();
        c->bnd.push_back(i);
        break;
      }
    }
  // Real harmonics at the outer-sphere points (Ylmpt, general.f90:11545-11568)
  // (one representation: only its harmonics, in increasing l*l + l + m)
  const int nall = (c->lmaxso + 1) * (c->lmaxso + 1);
  c->lmidx.clear();
  for (int l = 0; l <= c->lmaxso; l++)
    for (int m = -l; m <= l; m++)
      if (!sym || (is_c4_group(c->group) ? c4_harmonic_in(c->gmat, l, m, c->irrep)
                                         : harmonic_in(c->ops, harmonic_irrep(l, m), c->irrep)))
        c->lmidx.push_back(l * l + l + m);
  c->nlmso = (int)c->lmidx.size();
  if (c->nlmso == 0) return -1;
  c->ylm.assign(c->bnd.size() * c->nlmso, 0.0);
  if (c->ylm_comp) c->ycmp.assign(c->bnd.size() * c->nlmso, cd(0));
  std::vector<double> yall(nall);
  for (size_t b = 0; b < c->bnd.size(); b++) {
    const GridPoint &q = p[c->bnd[b]];
    real_ylm(c->lmaxso, q.ix * c->adimp, q.iy * c->adimp, q.iz * c->adimp, yall.data());
    for (int k = 0; k < c->nlmso; k++) {
      c->ylm[b * c->nlmso + k] = yall[c->lmidx[k]];
      if (c->ylm_comp) {
        const int l = lof(c->lmidx[k]), m = c->lmidx[k] - l * l - l;
        c->ycmp[b * c->nlmso + k] = yc(m, yall[l * l + l + m], yall[l * l + l - m]);
      }
    }
  }
  if (c->natome > 0) return build_atoms(c);
  // Rows and profile (newind, mat.f90:1812-1884)
  const int nl = np + c->nlmso;
  c->nligne = nl;
  c->lb1.resize(nl);
  c->lb2.resize(nl);
  for (int ii = 1; ii <= nl; ii++) c->lb1[ii - 1] = c->lb2[ii - 1] = ii;
  for (int i = 0; i < np; i++) {
    int &l1 = c->lb1[i], &l2 = c->lb2[i];
    for (int v = 0; v < 12; v++) {
      const int j = c->nbr[(size_t)i * 12 + v];
      if (j < 0) return -1;  // an interior point always has its 12 neighbours on the grid
      if (j < np) {
        l1 = std::min(l1, j + 1);
        l2 = std::max(l2, j + 1);
      } else {
        l2 = nl;
      }
    }
  }
  int lo = np + 1;
  for (int b = 0; b < c->nsortf; b++) lo = std::min(lo, c->bnd[b] + 1);
  for (int ii = np + 1; ii <= nl; ii++) {
    c->lb1[ii - 1] = lo;
    c->lb2[ii - 1] = nl;
  }
  for (int ii = nl - 1; ii >= 1; ii--) c->lb1[ii - 1] = std::min(c->lb1[ii - 1], c->lb1[ii]);
  // Compressed-row structure: diagonal, interior neighbours, outer columns
  // (grid rows touching the outside); outer rows: boundary points then the
  // outer columns.  Columns ascending.
  c->rp.assign(nl + 1, 0);
  c->cols.clear();
  std::vector<int> tmp;
  for (int i = 0; i < np; i++) {
    tmp.clear();
    tmp.push_back(i + 1);
    bool outside = false;
    for (int v = 0; v < 12; v++) {
      const int j = c->nbr[(size_t)i * 12 + v];
      if (j < np)
        tmp.push_back(j + 1);
      else
        outside = true;
    }
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());  // mirror images can coincide
    if (outside)
      for (int l = 0; l < c->nlmso; l++) tmp.push_back(np + l + 1);
    c->cols.insert(c->cols.end(), tmp.begin(), tmp.end());
    c->rp[i + 1] = (int64_t)c->cols.size();
  }
  for (int l = 0; l < c->nlmso; l++) {
    for (int b = 0; b < c->nsortf; b++) c->cols.push_back(c->bnd[b] + 1);
    for (int k = 0; k < c->nlmso; k++) c->cols.push_back(np + k + 1);
    c->rp[np + l + 1] = (int64_t)c->cols.size();
  }
  c->npr = np;
  c->nso = np;
  c->row.resize(np);
  for (int i = 0; i < np; i++) c->row[i] = i;
  return 0;
}


// stencil term of grid point i towards direction d: -clapl(v) times the
// character of the mirrors that bring the neighbour into the wedge, times
// rvol(i) (calcMatRow, mat.f90:1164-1222; -clapl(v) without symmetry)
inline double stencil(const fdmgpu_cluster *c, int i, int d) {
  return (character(c->irrep, c->flip[(size_t)i * 12 + d]) * -c->clapl[d + 1]) * c->rvol[i];
}

// Scalar-relativistic terms of grid point i (Relativiste, mat.f90:1155-1156,
// :1171-1185): with the potential's gradient G (npoint x 3, point-major),
// the diagonal loses a2s4 (V - E)^2 and the stencil term of direction d
// gains bder cgrad(v) dV/dx_k (k the axis of d), bder = -a2s4 / (1 - a2s4
// (V - E)), a2s4 = alpha^2 / 4; without G both are the plain terms.
struct RelPoint {
  double diag;  // V - E - clapl(0) (- a2s4 (V - E)^2)
  double bder;
};
inline RelPoint rel_point(const fdmgpu_cluster *c, double V, double Enervide, const double *G) {
  const double vme = V - Enervide;
  RelPoint r{vme - c->clapl[0], 0.0};
  if (G) {
    const double a2s4 = 0.25 * (kAlfaSf * kAlfaSf);
    r.diag = r.diag - a2s4 * (vme * vme);
    r.bder = -a2s4 / (1 - a2s4 * vme);
  }
  return r;
}
// stencil term with the relativistic gradient term (the plain term without G)
inline double stencil_rel(const fdmgpu_cluster *c, int i, int d, double bder, const double *G) {
  if (!G) return stencil(c, i, d);
  const double cf = -c->clapl[d + 1] + bder * c->cgrad[d + 1] * G[(size_t)i * 3 + (d >> 1) % 3];
  return (character(c->irrep, c->flip[(size_t)i * 12 + d]) * cf) * c->rvol[i];
}
// The same with the complex characters of 4, -4, 4/m: (Kar cf, Kari cf)
// rvol(i) (calcMatRow's Repres_comp branch, mat.f90:1217-1223)
inline cd stencil_c(const fdmgpu_cluster *c, int i, int d, double bder, const double *G) {
  if (!is_c4_group(c->group)) return cd(stencil_rel(c, i, d, bder, G), 0.0);
  double cf = -c->clapl[d + 1];
  if (G) cf = cf + bder * c->cgrad[d + 1] * G[(size_t)i * 3 + (d >> 1) % 3];
  const cd chi = c->gchar[c->gop[(size_t)i * 12 + d]];
  return cd((chi.real() * cf) * c->rvol[i], (chi.imag() * cf) * c->rvol[i]);
}
// the outer unknown k's harmonic at outer point b: Yc with complex harmonics
inline cd ysort(const fdmgpu_cluster *c, int b, int k) {
  return c->ylm_comp ? c->ycmp[(size_t)b * c->nlmso + k] : cd(c->ylm[(size_t)b * c->nlmso + k], 0.0);
}

// Outer-sphere functions at every boundary point (phiso, non-Rydberg branch,
// mat.f90:771-792, :975-994): k = sqrt(Ec + i Eimag), fnorm = sqrt(k/pi);
// B, N: (nb, lmaxso + 1)
void outer_functions(const fdmgpu_cluster *c, double Eimag, double Ecinetic_out, int E_comp, cd *B, cd *N) {
  const int nb = (int)c->bnd.size(), L = c->lmaxso;
  if (E_comp) {
    cd k = std::sqrt(cd(Ecinetic_out, Eimag));
    if (std::abs(k) < 1e-10) k = cd(1e-10, 0);
    const cd fn = std::sqrt(k / kPi);
    for (int b = 0; b < nb; b++)
      bessel_neumann<cd>(fn, k * c->rad[c->bnd[b]], L, &B[(size_t)b * (L + 1)], &N[(size_t)b * (L + 1)]);
  } else {
    const double k = std::sqrt(Ecinetic_out);
    const double fn = std::sqrt(k / kPi);
    std::vector<double> br(L + 1), nr(L + 1);
    for (int b = 0; b < nb; b++) {
      bessel_neumann<double>(fn, k * c->rad[c->bnd[b]], L, br.data(), nr.data());
      for (int l = 0; l <= L; l++) {
        B[(size_t)b * (L + 1) + l] = br[l];
        N[(size_t)b * (L + 1) + l] = nr[l];
      }
    }
  }
}

// Stored position of a 1-based column in row r (0-based) of the compressed rows
int64_t position(const fdmgpu_cluster *c, int r, int column) {
  return std::lower_bound(c->cols.begin() + c->rp[r], c->cols.begin() + c->rp[r + 1], column) - c->cols.begin();
}

// The energy-independent tables of a cluster with atoms, as
// fdmgpu_cluster_assembly_tables describes them, with rows in place of
// points: base holds the grid rows' -clapl(v) towards interstitial
// neighbours (mat.f90:1227-1232), the outer rows' -poidso Ylm (:1476-1483)
// and the atom rows' -poidsa Ylm (:1560-1581) in the border points' columns;
// a point inside an atom has no diagonal (qdiag -1).
int atom_cluster_tables(const fdmgpu_cluster *c, double *base, int64_t *qdiag, int *orow, int64_t *oq, int *optr,
                        int *ob, double *ow, double *ylm, double *brad, double *par) {
  const int np = c->npoint, nl = c->nlmso;
  std::fill(base, base + c->cols.size(), 0.0);
  int n = 0, k = 0;
  optr[0] = 0;
  for (int i = 0; i < np; i++) {
    if (c->numia[i]) {
      qdiag[i] = -1;
      continue;
    }
    const int r = c->row[i];
    qdiag[i] = position(c, r, r + 1);
    bool out = false;
    for (int d = 0; d < 12; d++) {
      const int j = c->nbr[(size_t)i * 12 + d];
      if (j < np && !c->numia[j]) {
        base[position(c, r, c->row[j] + 1)] += -c->clapl[d + 1];
      } else if (j >= np) {
        ob[k] = c->slot[j];
        ow[k] = -c->clapl[d + 1];
        k++;
        out = true;
      }
    }
    if (out) {
      orow[n] = r;
      oq[n] = c->irregular ? -1 : position(c, r, c->nso + 1);  // the irregular solution has no outer columns
      optr[++n] = k;
    }
  }
  for (int lm = 0; lm < nl && !c->irregular; lm++) {
    const int r = c->nso + lm;
    for (int b = 0; b < c->nsortf; b++)
      base[position(c, r, c->row[c->bnd[b]] + 1)] = -(c->wso[b] * c->ylm[(size_t)b * nl + lm]);
  }
  for (int a = 0; a < c->natome; a++)
    for (int lm = 0; lm < c->nri * c->nlmsa[a]; lm++) {  // irregular: each row block on its half of the border
      const int r = c->ianew[a] + lm, ri = lm >= c->nlmsa[a], l0 = lm - ri * c->nlmsa[a];
      const int nh = c->nbordf[a] / 2;
      for (int b = 0; b < c->nbordf[a]; b++) {
        if (c->irregular && (ri ? b < nh : b >= nh)) continue;
        base[position(c, r, c->row[c->abord[a][b]] + 1)] = -(c->apoids[a][b] * c->aylm[a][(size_t)b * c->nlmsa[a] + l0]);
      }
    }
  std::memcpy(ylm, c->ylm.data(), sizeof(double) * c->ylm.size());
  for (size_t b = 0; b < c->bnd.size(); b++) brad[b] = c->rad[c->bnd[b]];
  par[0] = c->clapl[0];
  par[1] = 4 * kPi / c->nsortf;
  return 0;
}

}  // namespace

extern "C" {

int fdmgpu_cluster_create_group(double rsort, double adimp, int lmaxso, int group, int irrep, fdmgpu_cluster **out) {
  return fdmgpu_cluster_create_group_ylm(rsort, adimp, lmaxso, group, irrep, is_c4_group(group) ? 1 : 0, out);
}

int fdmgpu_cluster_create_group_ylm(double rsort, double adimp, int lmaxso, int group, int irrep, int ylm_comp,
                                    fdmgpu_cluster **out) {
  // complex harmonics: the groups whose operations keep each Yc(l, m) (the
  // reference's subgroups accepting complex harmonics, tab_data.f90:547);
  // 4, -4 and 4/m always take them (Atom_comp_cal, general.f90:9337)
  if (ylm_comp && !(group <= 5 || is_c4_group(group))) return -1;
  if (!ylm_comp && is_c4_group(group)) return -1;
  if (is_c4_group(group)) {
    if (!out || !(rsort > 0) || !(adimp > 0) || lmaxso < 0 || lmaxso > 40 || irrep < 0 ||
        irrep > (group == 11 ? 7 : 3))
      return -1;
    fdmgpu_cluster *c = new (std::nothrow) fdmgpu_cluster();
    if (!c) return -3;
    c->rsort = rsort;
    c->adimp = adimp;
    c->lmaxso = lmaxso;
    c->irrep = irrep;
    c->group = group;
    c->ops = {0};
    c->ylm_comp = true;
    c->gmat = c4_ops(group);
    for (const auto &M : c->gmat) c->gchar.push_back(c4_char(M, irrep));
    if (build(c)) {
      delete c;
      return -1;
    }
    *out = c;
    return 0;
  }
  // the reference's point-group numbers (general.f90:5207-5220) of the
  // mirror groups, in posequiv's orientation: m's mirror is normal to z and
  // mm2's two mirrors are normal to x and y (general.f90:8961-8978)
  // The other subgroups of mmm (-1 = 2, 2 = 4 about z, 2/m = 5, 222 = 7)
  // have rotations or the inversion, with posequiv's areas: their
  // representations are again one-dimensional with characters +-1.
  int mirrors;
  switch (group) {
    case 1: mirrors = 0; break;
    case 3: mirrors = 4; break;
    case 6: mirrors = 3; break;
    case 8: mirrors = 7; break;
    case 2: case 4: case 5: case 7: mirrors = 7; break;  // irrep: a character vector over the axes
    default: return -1;
  }
  if (!out || !(rsort > 0) || !(adimp > 0) || lmaxso < 0 || lmaxso > 40 || irrep < -1 || (irrep & ~mirrors) > 0)
    return -1;
  const std::vector<int> ops = group_ops(group);
  if (irrep >= 0 && group != 1) {
    // irrep must be the smallest mask with its characters on the group
    for (int v = 0; v < irrep; v++) {
      bool same = true;
      for (int g : ops) same &= ((__builtin_popcount(g & v) ^ __builtin_popcount(g & irrep)) & 1) == 0;
      if (same) return -1;
    }
  }
  fdmgpu_cluster *c = new (std::nothrow) fdmgpu_cluster();
  if (!c) return -3;
  c->rsort = rsort;
  c->adimp = adimp;
  c->lmaxso = lmaxso;
  c->irrep = group != 1 ? irrep : -1;
  c->mirrors = mirrors;
  c->group = group;
  c->ops = ops;
  c->ylm_comp = ylm_comp != 0;
  if (build(c)) {
    delete c;
    return -1;
  }
  *out = c;
  return 0;
}

int fdmgpu_cluster_create_mmm(double rsort, double adimp, int lmaxso, int irrep, fdmgpu_cluster **out) {
  return fdmgpu_cluster_create_group(rsort, adimp, lmaxso, irrep < 0 ? 1 : 8, irrep, out);
}

int fdmgpu_cluster_create(double rsort, double adimp, int lmaxso, fdmgpu_cluster **out) {
  return fdmgpu_cluster_create_group(rsort, adimp, lmaxso, 1, -1, out);
}

int fdmgpu_cluster_create_atoms(double rsort, double adimp, int lmaxso, int natome, const double *posi,
                                const double *rmt, const int *lmax, fdmgpu_cluster **out) {
  return fdmgpu_cluster_create_atoms_ri(rsort, adimp, lmaxso, natome, posi, rmt, lmax, 0, out);
}

int fdmgpu_cluster_create_atoms_ri(double rsort, double adimp, int lmaxso, int natome, const double *posi,
                                   const double *rmt, const int *lmax, int irregular, fdmgpu_cluster **out) {
  if (!out || !(rsort > 0) || !(adimp > 0) || lmaxso < 0 || lmaxso > 40 || natome < 1 || natome > 32767 || !posi ||
      !rmt || !lmax || irregular < 0 || irregular > 1)
    return -1;
  double prev = 0;
  for (int a = 0; a < natome; a++) {
    const double d = std::sqrt(posi[3 * a] * posi[3 * a] + posi[3 * a + 1] * posi[3 * a + 1] +
                               posi[3 * a + 2] * posi[3 * a + 2]);
    if (!(rmt[a] > 0) || lmax[a] < 0 || lmax[a] > 12 || d < prev) return -1;  // sorted by distance, as posi
    prev = d;
  }
  fdmgpu_cluster *c = new (std::nothrow) fdmgpu_cluster();
  if (!c) return -3;
  c->rsort = rsort;
  c->adimp = adimp;
  c->lmaxso = lmaxso;
  c->irrep = -1;
  c->mirrors = 0;
  c->natome = natome;
  c->irregular = irregular != 0;
  c->nri = c->irregular ? 2 : 1;
  c->apos.assign(posi, posi + 3 * natome);
  c->armt.assign(rmt, rmt + natome);
  c->almax.assign(lmax, lmax + natome);
  if (build(c)) {
    delete c;
    return -1;
  }
  *out = c;
  return 0;
}

int fdmgpu_cluster_atom_sizes(const fdmgpu_cluster *c, int64_t *sizes) {
  int64_t nbord = 0, nphi = 0;
  for (int a = 0; a < c->natome; a++) {
    nbord += (int64_t)c->abord[a].size();
    nphi += (int64_t)c->abord[a].size() * c->nlmsa[a];
  }
  sizes[0] = c->natome;
  sizes[1] = c->npr;
  sizes[2] = nbord;
  sizes[3] = nphi;
  return 0;
}

int fdmgpu_cluster_atom(const fdmgpu_cluster *c, int a, int *info, int *ibord, double *poidsa, double *ylmato) {
  if (a < 0 || a >= c->natome) return -1;
  const int nb = (int)c->abord[a].size();
  info[0] = nb;
  info[1] = c->nbordf[a];
  info[2] = c->nlmsa[a];
  info[3] = c->ianew[a];
  if (ibord) std::memcpy(ibord, c->abord[a].data(), sizeof(int) * nb);
  if (poidsa) std::memcpy(poidsa, c->apoids[a].data(), sizeof(double) * c->nbordf[a]);
  if (ylmato) std::memcpy(ylmato, c->aylm[a].data(), sizeof(double) * c->aylm[a].size());
  return 0;
}

int fdmgpu_cluster_point_rows(const fdmgpu_cluster *c, int *numia, int *row) {
  for (int i = 0; i < c->npoint; i++) {
    if (numia) numia[i] = c->natome ? c->numia[i] : 0;
    if (row) row[i] = c->row[i];
  }
  return 0;
}

int fdmgpu_cluster_rvol(const fdmgpu_cluster *c, double *rvol) {
  std::memcpy(rvol, c->rvol.data(), sizeof(double) * c->npoint);
  return 0;
}

int fdmgpu_cluster_harmonics(const fdmgpu_cluster *c, int *lm) {
  std::memcpy(lm, c->lmidx.data(), sizeof(int) * c->nlmso);
  return 0;
}

void fdmgpu_cluster_destroy(fdmgpu_cluster *c) { delete c; }

int fdmgpu_cluster_sizes(const fdmgpu_cluster *c, int *s) {
  s[0] = c->npoint;
  s[1] = c->npso;
  s[2] = c->nligne;
  s[3] = c->nlmso;
  s[4] = c->nsortf;
  s[5] = (int)c->bnd.size();
  return 0;
}

int fdmgpu_cluster_points(const fdmgpu_cluster *c, double *xyz) {
  for (int i = 0; i < c->npoint; i++) {
    xyz[3 * i] = c->pt[i].ix * c->adimp;
    xyz[3 * i + 1] = c->pt[i].iy * c->adimp;
    xyz[3 * i + 2] = c->pt[i].iz * c->adimp;
  }
  return 0;
}

int fdmgpu_cluster_profile(const fdmgpu_cluster *c, int *lb1, int *lb2) {
  std::memcpy(lb1, c->lb1.data(), sizeof(int) * c->nligne);
  std::memcpy(lb2, c->lb2.data(), sizeof(int) * c->nligne);
  return 0;
}

int64_t fdmgpu_cluster_nnz(const fdmgpu_cluster *c) { return (int64_t)c->cols.size(); }

int fdmgpu_cluster_assemble(const fdmgpu_cluster *c, const double *V, double Enervide, double Eimag,
                            double Ecinetic_out, int E_comp, int basereel, int64_t *row_ptr, int *col, double *re,
                            double *im, double *rhs_re, double *rhs_im) {
  return fdmgpu_cluster_assemble_rel(c, V, nullptr, Enervide, Eimag, Ecinetic_out, E_comp, basereel, row_ptr, col, re,
                                     im, rhs_re, rhs_im);
}

int fdmgpu_cluster_assemble_rel(const fdmgpu_cluster *c, const double *V, const double *gradvr, double Enervide,
                                double Eimag, double Ecinetic_out, int E_comp, int basereel, int64_t *row_ptr,
                                int *col, double *re, double *im, double *rhs_re, double *rhs_im) {
  if (c->natome > 0) return -2;  // fdmgpu_cluster_assemble_atoms
  const double *G = gradvr;
  // complex harmonics or characters make the arithmetic complex (mat.f90:97-104)
  const bool cx = c->ylm_comp || is_c4_group(c->group);
  const bool cal = E_comp || !basereel || cx;
  if (cx && !im) return -1;
  const int np = c->npoint, nl = c->nlmso, nb = (int)c->bnd.size(), L = c->lmaxso;
  std::vector<cd> B((size_t)nb * (L + 1)), N((size_t)nb * (L + 1));
  outer_functions(c, Eimag, Ecinetic_out, E_comp, B.data(), N.data());
  // Outer-column factor of an outside point (calcMatRow, mat.f90:1330-1346):
  // the Neumann function in the real basis, j + i n in the complex one.
  auto cfac = [&](int b, int l) -> cd {
    const cd bb = B[(size_t)b * (L + 1) + l], nn = N[(size_t)b * (L + 1) + l];
    if (basereel) return nn;
    return bb + cd(0, 1) * nn;
  };
  std::memcpy(row_ptr, c->rp.data(), sizeof(int64_t) * (c->nligne + 1));
  std::memcpy(col, c->cols.data(), sizeof(int) * c->cols.size());
  if (rhs_re) std::memset(rhs_re, 0, sizeof(double) * (size_t)c->nligne * nl);
  if (rhs_im && cal) std::memset(rhs_im, 0, sizeof(double) * (size_t)c->nligne * nl);
  std::vector<cd> row(nl), rrow(nl);
  for (int i = 0; i < np; i++) {
    const int64_t q0 = c->rp[i], q1 = c->rp[i + 1];
    std::fill(row.begin(), row.end(), cd(0));
    std::fill(rrow.begin(), rrow.end(), cd(0));
    const RelPoint rp = rel_point(c, V[i], Enervide, G);
    const cd diag(rp.diag, E_comp ? -Eimag : 0.0);  // mat.f90:1147-1150, :1155
    // stencil term of direction d: -clapl(v) times the character of the
    // mirrors that bring the neighbour into the wedge, times rvol(i)
    // (mat.f90:1164-1222; 1 and 1 without symmetry)
    // interior entries in column order, each neighbour's terms over rvol(j)
    // added in direction order (:1225-1231); a neighbour that is the point's
    // own mirror image adds to the diagonal
    int64_t q = q0;
    for (; q < q1 && col[q] <= np; q++) {
      const int j = col[q] - 1;
      cd v = (j == i) ? diag : cd(0);
      for (int d = 0; d < 12; d++)
        if (c->nbr[(size_t)i * 12 + d] == j) {
          if (cx)
            v += stencil_c(c, i, d, rp.bder, G) / c->rvol[j];
          else
            v += stencil_rel(c, i, d, rp.bder, G) / c->rvol[j];
        }
      re[q] = v.real();
      if (cal) im[q] = v.imag();
    }
    for (int d = 0; d < 12; d++) {  // outside neighbours couple to the outer unknowns, mat.f90:1312-1398
      const int j = c->nbr[(size_t)i * 12 + d];
      if (j < np) continue;
      const int b = c->slot[j];
      if (cx) {  // complex harmonics (Ycomp) and characters, :1328-1377
        const cd w = stencil_c(c, i, d, rp.bder, G);
        for (int lm = 0; lm < nl; lm++) {
          const int l = lof(c->lmidx[lm]);
          const cd y = ysort(c, b, lm);
          row[lm] += cfac(b, l) * y * w;
          rrow[lm] -= y * B[(size_t)b * (L + 1) + l] * w;
        }
        continue;
      }
      const double w = stencil_rel(c, i, d, rp.bder, G);
      for (int lm = 0; lm < nl; lm++) {
        const int l = lof(c->lmidx[lm]);
        const double y = c->ylm[(size_t)b * nl + lm];
        row[lm] += cfac(b, l) * y * w;
        rrow[lm] -= y * B[(size_t)b * (L + 1) + l] * w;
      }
    }
    for (int lm = 0; q < q1; q++, lm++) {
      re[q] = row[lm].real();
      if (cal) im[q] = row[lm].imag();
    }
    if (rhs_re)
      for (int lm = 0; lm < nl; lm++) {
        rhs_re[(size_t)i * nl + lm] = rrow[lm].real();
        if (cal && rhs_im) rhs_im[(size_t)i * nl + lm] = rrow[lm].imag();
      }
  }
  for (int lm = 0; lm < nl && c->ylm_comp; lm++) {  // outer rows with complex harmonics, :1425-1497
    const int ii = np + lm;
    int64_t q = c->rp[ii];
    for (int b = 0; b < c->nsortf; b++, q++) {
      const cd f = c->wso[b] * std::conj(ysort(c, b, lm));
      re[q] = -f.real() / c->rvol[c->bnd[b]];
      im[q] = -f.imag() / c->rvol[c->bnd[b]];
    }
    for (int lp = 0; lp < nl; lp++, q++) {
      const int l = lof(c->lmidx[lp]);
      cd s = 0;
      for (int b = 0; b < c->nsortf; b++) s += (c->wso[b] * (std::conj(ysort(c, b, lm)) * ysort(c, b, lp))) * cfac(b, l);
      re[q] = s.real();
      im[q] = s.imag();
    }
    if (rhs_re)
      for (int lf = 0; lf < nl; lf++) {
        const int l = lof(c->lmidx[lf]);
        cd s = 0;
        for (int b = 0; b < c->nsortf; b++)
          s += ((c->wso[b] * std::conj(ysort(c, b, lm))) * ysort(c, b, lf)) * B[(size_t)b * (L + 1) + l];
        rhs_re[(size_t)ii * nl + lf] = -s.real();
        if (rhs_im) rhs_im[(size_t)ii * nl + lf] = -s.imag();
      }
  }
  for (int lm = 0; lm < nl && !c->ylm_comp; lm++) {  // outer rows, mat.f90:1401-1511, weights poidso (bordure)
    const int ii = np + lm;
    int64_t q = c->rp[ii];
    for (int b = 0; b < c->nsortf; b++, q++) {  // :1476-1483
      re[q] = -(c->wso[b] * c->ylm[(size_t)b * nl + lm]) / c->rvol[c->bnd[b]];
      if (cal) im[q] = 0;
    }
    for (int lp = 0; lp < nl; lp++, q++) {
      const int l = lof(c->lmidx[lp]);
      cd s = 0;
      for (int b = 0; b < c->nsortf; b++)
        s += (c->wso[b] * c->ylm[(size_t)b * nl + lm] * c->ylm[(size_t)b * nl + lp]) * cfac(b, l);
      re[q] = s.real();
      if (cal) im[q] = s.imag();
    }
    if (rhs_re)
      for (int lf = 0; lf < nl; lf++) {
        const int l = lof(c->lmidx[lf]);
        cd s = 0;
        for (int b = 0; b < c->nsortf; b++)
          s += (c->wso[b] * c->ylm[(size_t)b * nl + lm] * c->ylm[(size_t)b * nl + lf]) * B[(size_t)b * (L + 1) + l];
        rhs_re[(size_t)ii * nl + lf] = -s.real();
        if (cal && rhs_im) rhs_im[(size_t)ii * nl + lf] = -s.imag();
      }
  }
  return 0;
}

int fdmgpu_cluster_assemble_atoms(const fdmgpu_cluster *c, const double *V, double Enervide, double Eimag,
                                  double Ecinetic_out, int E_comp, int basereel, const double *phi_re,
                                  const double *phi_im, int64_t *row_ptr, int *col, double *re, double *im,
                                  double *rhs_re, double *rhs_im) {
  return fdmgpu_cluster_assemble_atoms_ri(c, V, Enervide, Eimag, Ecinetic_out, E_comp, basereel, phi_re, phi_im,
                                          nullptr, nullptr, row_ptr, col, re, im, rhs_re, rhs_im);
}

int fdmgpu_cluster_assemble_atoms_ri(const fdmgpu_cluster *c, const double *V, double Enervide, double Eimag,
                                     double Ecinetic_out, int E_comp, int basereel, const double *phi_re,
                                     const double *phi_im, const double *phi2_re, const double *phi2_im,
                                     int64_t *row_ptr, int *col, double *re, double *im, double *rhs_re,
                                     double *rhs_im) {
  return fdmgpu_cluster_assemble_atoms_rel(c, V, nullptr, Enervide, Eimag, Ecinetic_out, E_comp, basereel, phi_re,
                                           phi_im, phi2_re, phi2_im, row_ptr, col, re, im, rhs_re, rhs_im);
}

int fdmgpu_cluster_assemble_atoms_rel(const fdmgpu_cluster *c, const double *V, const double *gradvr,
                                      double Enervide, double Eimag, double Ecinetic_out, int E_comp, int basereel,
                                      const double *phi_re, const double *phi_im, const double *phi2_re,
                                      const double *phi2_im, int64_t *row_ptr, int *col, double *re, double *im,
                                      double *rhs_re, double *rhs_im) {
  const double *G = gradvr;
  const bool cal = E_comp || !basereel;
  if (!c || !V || !phi_re || !row_ptr || !col || !re || (cal && !im) || (phi_im && !cal)) return -1;
  // the irregular solution exists with complex energy only (FDM_comp), and
  // needs the irregular atomic solutions; a regular cluster takes none
  if (c->natome == 0 || (c->irregular && (!E_comp || basereel || !phi2_re || (phi_im && !phi2_im))) ||
      (!c->irregular && phi2_re))
    return -1;
  const int np = c->npoint, nl = c->nlmso, nb = (int)c->bnd.size(), L = c->lmaxso, nlig = c->nligne;
  std::vector<cd> B((size_t)nb * (L + 1)), N((size_t)nb * (L + 1));
  outer_functions(c, Eimag, Ecinetic_out, E_comp, B.data(), N.data());
  auto cfac = [&](int b, int l) -> cd {  // mat.f90:1330-1346
    const cd bb = B[(size_t)b * (L + 1) + l], nn = N[(size_t)b * (L + 1) + l];
    if (basereel) return nn;
    return bb + cd(0, 1) * nn;
  };
  std::vector<size_t> phoff(c->natome + 1, 0);  // each atom's (nbord, nlmsa) block of phiato
  for (int a = 0; a < c->natome; a++) phoff[a + 1] = phoff[a] + c->abord[a].size() * c->nlmsa[a];
  std::memcpy(row_ptr, c->rp.data(), sizeof(int64_t) * (nlig + 1));
  std::memcpy(col, c->cols.data(), sizeof(int) * c->cols.size());
  std::fill(re, re + c->cols.size(), 0.0);
  if (cal) std::fill(im, im + c->cols.size(), 0.0);
  if (rhs_re) std::memset(rhs_re, 0, sizeof(double) * (size_t)nlig * nl);
  if (rhs_im && cal) std::memset(rhs_im, 0, sizeof(double) * (size_t)nlig * nl);
  int64_t q0 = 0, q1 = 0;
  auto at = [&](int column) -> int64_t {  // stored position of a 1-based column in the current row
    return std::lower_bound(c->cols.begin() + q0, c->cols.begin() + q1, column) - c->cols.begin();
  };
  auto add = [&](int column, double vr, double vi) {
    const int64_t q = at(column);
    re[q] += vr;
    if (cal) im[q] += vi;
  };
  for (int r = 0; r < nlig; r++) {
    q0 = c->rp[r];
    q1 = c->rp[r + 1];
    if (c->rkind[r] == 0) {  // grid row, mat.f90:1142-1398 (P1: rvol = 1, no character)
      const int i = c->ridx[r];
      const RelPoint rp = rel_point(c, V[i], Enervide, G);
      add(r + 1, rp.diag, E_comp ? -Eimag : 0.0);
      for (int d = 0; d < 12; d++) {
        const int j = c->nbr[(size_t)i * 12 + d];
        const double w = stencil_rel(c, i, d, rp.bder, G);
        if (j < np && !c->numia[j]) {  // :1227-1232
          add(c->row[j] + 1, w, 0.0);
        } else if (j < np) {  // a neighbour inside atom a couples to its expansion, :1233-1309
          const int a = c->numia[j] - 1;
          const std::vector<int> &bl = c->abord[a];
          const int ib = (int)(std::find(bl.begin() + c->nbordf[a], bl.end(), j) - bl.begin());
          for (int ri = 0; ri < c->nri; ri++) {  // regular, then the irregular atomic solutions (:1266-1307)
            const double *fr = ri ? phi2_re : phi_re, *fi = ri ? phi2_im : phi_im;
            const double *pr = fr + phoff[a] + (size_t)ib * c->nlmsa[a];
            const double *pi = fi ? fi + phoff[a] + (size_t)ib * c->nlmsa[a] : nullptr;
            for (int lm = 0; lm < c->nlmsa[a]; lm++)
              add(c->ianew[a] + ri * c->nlmsa[a] + lm + 1, w * pr[lm], pi ? w * pi[lm] : 0.0);
          }
        } else if (c->irregular) {  // no scattered wave; incoming -i j + n, :1318-1398
          const int b = c->slot[j];
          if (rhs_re)
            for (int lm = 0; lm < nl; lm++) {
              const int l = lof(c->lmidx[lm]);
              const double y = c->ylm[(size_t)b * nl + lm];
              const cd s = y * (cd(0, -1) * B[(size_t)b * (L + 1) + l] + N[(size_t)b * (L + 1) + l]) * w;
              rhs_re[(size_t)r * nl + lm] -= s.real();
              if (rhs_im) rhs_im[(size_t)r * nl + lm] -= s.imag();
            }
        } else {  // outside the sphere: the outer unknowns and the incoming wave, :1312-1398
          const int b = c->slot[j];
          for (int lm = 0; lm < nl; lm++) {
            const int l = lof(c->lmidx[lm]);
            const double y = c->ylm[(size_t)b * nl + lm];
            const cd v = cal ? cfac(b, l) * y * w : cd(w * (y * N[(size_t)b * (L + 1) + l].real()));
            add(c->nso + lm + 1, v.real(), v.imag());
            if (rhs_re) {
              const cd s = cal ? y * B[(size_t)b * (L + 1) + l] * w : cd(w * (y * B[(size_t)b * (L + 1) + l].real()));
              rhs_re[(size_t)r * nl + lm] -= s.real();
              if (cal && rhs_im) rhs_im[(size_t)r * nl + lm] -= s.imag();
            }
          }
        }
      }
    } else if (c->rkind[r] == 1) {  // expansion row Lm of atom a, mat.f90:1515-1632
      const int a = c->ridx[r] >> 16, na = c->nlmsa[a];
      const int ri = (c->ridx[r] & 0xffff) >= na;  // the irregular solution's second row block
      const int lm = (c->ridx[r] & 0xffff) - ri * na;
      const int nh = c->nbordf[a] / 2;  // irregular: the blocks match on the two halves of the border
      const std::vector<double> &pw = c->apoids[a], &Y = c->aylm[a];
      for (int b = 0; b < c->nbordf[a]; b++) {
        if (c->irregular && (ri ? b < nh : b >= nh)) continue;
        add(c->row[c->abord[a][b]] + 1, -(pw[b] * Y[(size_t)b * na + lm]), 0.0);
      }
      for (int rj = 0; rj < c->nri; rj++) {  // :1580-1630, the columns of the regular then irregular block
        const double *fr = rj ? phi2_re : phi_re, *fi = rj ? phi2_im : phi_im;
        const int b0 = c->irregular && rj ? nh : 0, b1 = c->irregular && !rj ? nh : c->nbordf[a];
        for (int lp = 0; lp < na; lp++) {
          double sr = 0, si = 0;
          for (int b = b0; b < b1; b++) {
            const double y = Y[(size_t)b * na + lm], pr = fr[phoff[a] + (size_t)b * na + lp];
            if (cal) {
              sr += pw[b] * (pr * y);
              if (fi) si += pw[b] * (fi[phoff[a] + (size_t)b * na + lp] * y);
            } else {
              sr += pw[b] * y * pr;
            }
          }
          add(c->ianew[a] + rj * na + lp + 1, sr, si);
        }
      }
    } else {  // outer row lm, mat.f90:1401-1511
      const int lm = c->ridx[r];
      for (int lp = 0; lp < nl; lp++) {
        const int l = lof(c->lmidx[lp]);
        cd s = 0;
        for (int b = 0; b < c->nsortf; b++) {
          const double ym = c->ylm[(size_t)b * nl + lm], yp = c->ylm[(size_t)b * nl + lp];
          s += cal ? (c->wso[b] * (ym * yp)) * cfac(b, l) : cd(c->wso[b] * ym * yp * N[(size_t)b * (L + 1) + l].real());
        }
        add(c->nso + lp + 1, s.real(), s.imag());
      }
      for (int b = 0; b < c->nsortf; b++) {
        const double f = c->wso[b] * c->ylm[(size_t)b * nl + lm];
        add(c->row[c->bnd[b]] + 1, -f, 0.0);
        if (rhs_re)
          for (int lf = 0; lf < nl; lf++) {
            const int l = lof(c->lmidx[lf]);
            const cd s = (f * c->ylm[(size_t)b * nl + lf]) * B[(size_t)b * (L + 1) + l];
            rhs_re[(size_t)r * nl + lf] -= s.real();
            if (cal && rhs_im) rhs_im[(size_t)r * nl + lf] -= s.imag();
          }
      }
    }
  }
  return 0;
}


int fdmgpu_cluster_assembly_sizes(const fdmgpu_cluster *c, int64_t *sizes) {
  int64_t nrow = 0, nlink = 0;
  for (int i = 0; i < c->npoint; i++) {
    if (c->numia[i]) continue;
    int k = 0;
    for (int d = 0; d < 12; d++) k += c->nbr[(size_t)i * 12 + d] >= c->npoint;
    nrow += k > 0;
    nlink += k;
  }
  sizes[0] = nrow;
  sizes[1] = nlink;
  sizes[2] = (int64_t)c->bnd.size();
  sizes[3] = (int64_t)c->cols.size();
  return 0;
}

// The device assembler's tables of a cluster without atoms; with complex
// harmonics or characters (cx) the imaginary parts go to base_im, ow_im,
// ylm_im and the point's own images' terms to dself / dself_im.
static int cluster_tables(const fdmgpu_cluster *c, double *base, double *base_im, int64_t *qdiag, int *orow,
                          int64_t *oq, int *optr, int *ob, double *ow, double *ow_im, double *ylm, double *ylm_im,
                          double *brad, double *par, double *dself, double *dself_im) {
  const bool cx = c->ylm_comp || is_c4_group(c->group);
  const int np = c->npoint, nl = c->nlmso;
  std::fill(base, base + c->cols.size(), 0.0);
  if (cx) {
    std::fill(base_im, base_im + c->cols.size(), 0.0);
    std::fill(dself, dself + 3 * (size_t)np, 0.0);
    std::fill(dself_im, dself_im + 3 * (size_t)np, 0.0);
  }
  int r = 0, k = 0;
  optr[0] = 0;
  for (int i = 0; i < np; i++) {
    const int64_t q0 = c->rp[i], q1 = c->rp[i + 1];
    int64_t q = q0;
    for (; q < q1 && c->cols[q] <= np; q++) {
      const int j = c->cols[q] - 1;
      if (j == i) {  // E-dependent; a point's own images are in dself
        qdiag[i] = q;
        if (cx)
          for (int d = 0; d < 12; d++)
            if (c->nbr[(size_t)i * 12 + d] == i) {
              const cd v = stencil_c(c, i, d, 0.0, nullptr) / c->rvol[i];
              dself[3 * (size_t)i + (d % 6) / 2] += v.real();
              dself_im[3 * (size_t)i + (d % 6) / 2] += v.imag();
            }
        continue;
      }
      for (int d = 0; d < 12; d++)
        if (c->nbr[(size_t)i * 12 + d] == j) {
          if (cx) {
            const cd v = stencil_c(c, i, d, 0.0, nullptr) / c->rvol[j];
            base[q] += v.real();
            base_im[q] += v.imag();
          } else {
            base[q] += stencil(c, i, d) / c->rvol[j];
          }
        }
    }
    bool out = false;
    for (int d = 0; d < 12; d++) {
      const int j = c->nbr[(size_t)i * 12 + d];
      if (j < np) continue;
      ob[k] = c->slot[j];
      if (cx) {
        const cd w = stencil_c(c, i, d, 0.0, nullptr);
        ow[k] = w.real();
        ow_im[k] = w.imag();
      } else {
        ow[k] = stencil(c, i, d);
      }
      k++;
      out = true;
    }
    if (out) {
      orow[r] = i;
      oq[r] = q;
      optr[++r] = k;
    }
  }
  const double pw = 4 * kPi / c->nsortf;
  for (int lm = 0; lm < nl; lm++) {
    const int64_t q = c->rp[np + lm];
    for (int b = 0; b < c->nsortf; b++) {
      if (cx) {  // -poidso conj(Yc) / rvol (mat.f90:1476-1483 with Ylm_comp)
        const cd f = c->wso[b] * std::conj(ysort(c, b, lm));
        base[q + b] = -f.real() / c->rvol[c->bnd[b]];
        base_im[q + b] = -f.imag() / c->rvol[c->bnd[b]];
      } else {
        base[q + b] = -(c->wso[b] * c->ylm[(size_t)b * nl + lm]) / c->rvol[c->bnd[b]];
      }
    }
  }
  for (size_t b = 0; b < c->bnd.size(); b++)
    for (int lm = 0; lm < nl; lm++) {
      const cd y = ysort(c, (int)b, lm);
      ylm[b * nl + lm] = y.real();
      if (cx) ylm_im[b * nl + lm] = y.imag();
    }
  for (size_t b = 0; b < c->bnd.size(); b++) brad[b] = c->rad[c->bnd[b]];
  par[0] = c->clapl[0];
  par[1] = pw;
  return 0;
}

int fdmgpu_cluster_assembly_tables(const fdmgpu_cluster *c, double *base, int64_t *qdiag, int *orow,
                                   int64_t *oq, int *optr, int *ob, double *ow, double *ylm, double *brad,
                                   double *par) {
  // real harmonics and characters only (fdmgpu_cluster_assembly_tables_complex)
  if (c->ylm_comp || is_c4_group(c->group)) return -2;
  if (c->natome > 0) return atom_cluster_tables(c, base, qdiag, orow, oq, optr, ob, ow, ylm, brad, par);
  return cluster_tables(c, base, nullptr, qdiag, orow, oq, optr, ob, ow, nullptr, ylm, nullptr, brad, par, nullptr,
                        nullptr);
}

int fdmgpu_cluster_assembly_tables_complex(const fdmgpu_cluster *c, double *base, double *base_im, int64_t *qdiag,
                                           int *orow, int64_t *oq, int *optr, int *ob, double *ow, double *ow_im,
                                           double *ylm, double *ylm_im, double *brad, double *par, double *dself,
                                           double *dself_im) {
  if (!c || !base || !base_im || !qdiag || !orow || !oq || !optr || !ob || !ow || !ow_im || !ylm || !ylm_im ||
      !brad || !par || !dself || !dself_im)
    return -1;
  if (!(c->ylm_comp || is_c4_group(c->group)) || c->natome > 0 || c->irregular) return -2;
  return cluster_tables(c, base, base_im, qdiag, orow, oq, optr, ob, ow, ow_im, ylm, ylm_im, brad, par, dself,
                        dself_im);
}

int fdmgpu_cluster_assembly_weights(const fdmgpu_cluster *c, int *lharm, double *wso, double *dself) {
  for (int k = 0; k < c->nlmso; k++) lharm[k] = lof(c->lmidx[k]);
  std::memcpy(wso, c->wso.data(), sizeof(double) * c->nsortf);
  if (is_c4_group(c->group)) {  // complex characters: no real self terms (and no device tables)
    std::fill(dself, dself + 3 * (size_t)c->npoint, 0.0);
    return 0;
  }
  for (int i = 0; i < c->npoint; i++) {
    double *s = dself + 3 * (size_t)i;
    s[0] = s[1] = s[2] = 0.0;
    for (int d = 0; d < 12; d++)  // only the second-shell steps towards a mirror can land on the point
      if (c->nbr[(size_t)i * 12 + d] == i) s[(d % 6) / 2] += stencil(c, i, d) / c->rvol[i];
  }
  return 0;
}

int fdmgpu_cluster_rel_tables(const fdmgpu_cluster *c, int64_t *rq, double *rfac, double *rdiv, int *opoint,
                              int *odir, double *ofac, int *apoint, int *adir, double *clapl, double *cgrad) {
  if (!c || !rq || !rfac || !rdiv || !clapl || !cgrad) return -1;
  if (c->irregular || c->ylm_comp || is_c4_group(c->group)) return -2;  // regular, real harmonics and characters
  const int np = c->npoint;
  int r = 0, k = 0, n = 0, ka = 0;
  for (int i = 0; i < np; i++) {
    for (int d = 0; d < 12; d++) {
      rq[12 * (size_t)i + d] = -1;
      rfac[12 * (size_t)i + d] = 0;
      rdiv[12 * (size_t)i + d] = 1;
    }
    if (c->numia[i]) continue;
    const double chi_r = c->rvol[i];
    bool out = false;
    for (int d = 0; d < 12; d++) {  // the grid row's own entries, as in the assemblers
      const int j = c->nbr[(size_t)i * 12 + d];
      const double f = character(c->irrep, c->flip[(size_t)i * 12 + d]) * chi_r;
      if (j < np && !c->numia[j]) {
        rq[12 * (size_t)i + d] = position(c, c->row[i], c->row[j] + 1);
        rfac[12 * (size_t)i + d] = f;
        rdiv[12 * (size_t)i + d] = c->natome ? 1.0 : c->rvol[j];
      } else if (j >= np) {  // the outer-column links, in the assembly tables' order
        if (opoint && odir && ofac) {
          odir[k] = d;
          ofac[k] = f;
        }
        k++;
        out = true;
      }
    }
    if (out && opoint) opoint[r] = i;
    r += out;
    if (c->natome == 0) continue;
    std::vector<int> atoms;  // the atom couplings, in the atom tables' order
    for (int d = 0; d < 12; d++) {
      const int j = c->nbr[(size_t)i * 12 + d];
      if (j < np && c->numia[j] && std::find(atoms.begin(), atoms.end(), c->numia[j]) == atoms.end())
        atoms.push_back(c->numia[j]);
    }
    for (int a : atoms) {
      for (int d = 0; d < 12; d++) {
        const int j = c->nbr[(size_t)i * 12 + d];
        if (j < np && c->numia[j] == a) {
          if (adir) adir[ka] = d;
          ka++;
        }
      }
      if (apoint) apoint[n] = i;
      n++;
    }
  }
  for (int v = 0; v < 13; v++) {
    clapl[v] = c->clapl[v];
    cgrad[v] = c->cgrad[v];
  }
  return 0;
}

int fdmgpu_cluster_atom_assembly_sizes(const fdmgpu_cluster *c, int64_t *sizes) {
  int64_t nar = 0, nal = 0, nphi = 0, nrow = 0, nbf = 0, nyl = 0;
  for (int i = 0; i < c->npoint; i++) {
    if (c->numia[i]) continue;
    std::vector<int> seen;
    for (int d = 0; d < 12; d++) {
      const int j = c->nbr[(size_t)i * 12 + d];
      if (j >= c->npoint || !c->numia[j]) continue;
      nal++;
      if (std::find(seen.begin(), seen.end(), c->numia[j]) == seen.end()) seen.push_back(c->numia[j]);
    }
    nar += (int64_t)seen.size();
  }
  for (int a = 0; a < c->natome; a++) {
    nphi += (int64_t)c->abord[a].size() * c->nlmsa[a];
    nrow += c->nri * c->nlmsa[a];
    nbf += c->nbordf[a];
    nyl += (int64_t)c->nbordf[a] * c->nlmsa[a];
  }
  sizes[0] = c->natome;
  sizes[1] = nar;
  sizes[2] = nal;
  sizes[3] = nphi;
  sizes[4] = nrow;
  sizes[5] = nbf;
  sizes[6] = nyl;
  return 0;
}

int fdmgpu_cluster_atom_assembly_tables(const fdmgpu_cluster *c, int64_t *aq, int *aa, int *aptr, int64_t *alphi,
                                        double *alw, int *ianew, int *nlmsa, int *nbordf, int64_t *phoff,
                                        int64_t *bq, double *poidsa, double *ylmato) {
  if (c->natome < 1) return -1;
  std::vector<int64_t> off(c->natome + 1, 0);
  for (int a = 0; a < c->natome; a++) off[a + 1] = off[a] + (int64_t)c->abord[a].size() * c->nlmsa[a];
  int n = 0, k = 0;
  aptr[0] = 0;
  for (int i = 0; i < c->npoint; i++) {
    if (c->numia[i]) continue;
    std::vector<int> atoms;  // in the order the stencil first meets them
    for (int d = 0; d < 12; d++) {
      const int j = c->nbr[(size_t)i * 12 + d];
      if (j < c->npoint && c->numia[j] &&
          std::find(atoms.begin(), atoms.end(), c->numia[j] - 1) == atoms.end())
        atoms.push_back(c->numia[j] - 1);
    }
    for (int a : atoms) {
      for (int d = 0; d < 12; d++) {
        const int j = c->nbr[(size_t)i * 12 + d];
        if (j >= c->npoint || c->numia[j] != a + 1) continue;
        const std::vector<int> &bl = c->abord[a];
        const int ib = (int)(std::find(bl.begin() + c->nbordf[a], bl.end(), j) - bl.begin());
        alphi[k] = off[a] + (int64_t)ib * c->nlmsa[a];
        alw[k] = -c->clapl[d + 1];
        k++;
      }
      aq[n] = position(c, c->row[i], c->ianew[a] + 1);
      aa[n] = a;
      aptr[++n] = k;
    }
  }
  int64_t r = 0, pb = 0, yb = 0;
  for (int a = 0; a < c->natome; a++) {
    ianew[a] = c->ianew[a];
    nlmsa[a] = c->nlmsa[a];
    nbordf[a] = c->nbordf[a];
    phoff[a] = off[a];
    for (int lm = 0; lm < c->nri * c->nlmsa[a]; lm++) bq[r++] = position(c, c->ianew[a] + lm, c->ianew[a] + 1);
    for (int b = 0; b < c->nbordf[a]; b++) {
      poidsa[pb++] = c->apoids[a][b];
      for (int lm = 0; lm < c->nlmsa[a]; lm++) ylmato[yb++] = c->aylm[a][(size_t)b * c->nlmsa[a] + lm];
    }
  }
  return 0;
}

}  // extern "C"
