/*
 * fdmgpu_cluster — builder of the synthetic FDM systems the solver is
 * measured on: a cubic grid inside a sphere, ordered by radius, with the
 * 4th-order Laplacian stencil and the outer-sphere expansion rows, i.e. the
 * inputs the reference's mat_solve receives for a cluster without atoms in
 * point group P1 or a mirror group, or with muffin-tin atoms in P1
 * (fdmgpu_cluster_create_atoms).  It replaces the reference's set-up chain
 *   reseau   general.f90:9439   (grid and radial order)
 *   laplac   general.f90:9732   (stencil, neighbour table)
 *   bordure  general.f90:9895   (outer-sphere point lists and weights)
 *   Ylmpt    general.f90:11505  (real harmonics at the outer points)
 *   phiso    mat.f90:739        (Bessel/Neumann functions, non-Rydberg)
 *   newind   mat.f90:1761       (row numbering, profile lb1/lb2)
 *   calcMatRow mat.f90:1034     (rows of the matrix and incoming-wave RHS)
 * All lengths in bohr, energies in Rydberg.  Rows are 1..nligne: grid point
 * i (1-based, radial order) is row i, the nlmso outer-sphere unknowns are the
 * last rows.  Arithmetic is complex (cal_comp) when E_comp or !basereel
 * (mat.f90:96-104).
 */
#ifndef FDMGPU_CLUSTER_H
#define FDMGPU_CLUSTER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fdmgpu_cluster fdmgpu_cluster;

/* Sphere radius rsort and grid step adimp (bohr), outer expansion lmaxso
 * (nlmso = (lmaxso+1)^2 outer rows). */
int fdmgpu_cluster_create(double rsort, double adimp, int lmaxso, fdmgpu_cluster **out);
/* One representation of a cluster with the symmetry of the three mirror
 * planes x = 0, y = 0, z = 0 (point group mmm; the potential must have it):
 * the system of the igrph loop (fdm.f90:3775) for that representation.
 * irrep bits 0, 1, 2 set = odd under the mirror normal to x, y, z (-1 = the
 * whole cluster, as fdmgpu_cluster_create).  The grid is the wedge
 * x, y, z >= 0 (reseau with posequiv, general.f90:9512, :8892), each point
 * weighted by rvol = its number of mirror images / 8 (general.f90:9582-9651);
 * a neighbour outside the wedge is its mirror image with the
 * representation's character (laplac's isvois, general.f90:9839-9871;
 * calcMatRow, mat.f90:1168-1231); the outer-sphere weights are
 * 4 pi rvol / sum(rvol) (bordure, general.f90:10073-10075) and the outer
 * unknowns are the representation's harmonics in increasing l*l + l + m.
 * The unknown of grid row i is rvol(i) times the wave function there
 * (mat.f90:2132-2143).  fdmgpu_cluster_assembly_tables and
 * fdmgpu_cluster_assembly_weights give the device assembler's tables. */
int fdmgpu_cluster_create_mmm(double rsort, double adimp, int lmaxso, int irrep, fdmgpu_cluster **out);
/* The same for every subgroup of mmm by the reference's point-group number
 * (general.f90:5207-5220), in posequiv's orientation and with its
 * calculation areas (general.f90:8950-8978): group 1 (P1, irrep ignored),
 * 2 (-1: z > 0, then the plane z = 0 split at x > 0 and the half line),
 * 3 (m, the mirror z = 0: z >= 0), 4 (2, the two-fold axis z: x > 0, then
 * the half plane x = 0, y >= 0), 5 (2/m: the area of 2 with z >= 0), 6 (mm2,
 * the mirrors x = 0 and y = 0: x, y >= 0), 7 (222: x, y > 0, then the
 * quarter planes x = 0 and y = 0 with z >= 0) and 8 (mmm, as
 * fdmgpu_cluster_create_mmm).  Every operation negates some coordinates
 * (opsym: 2-fold axes 22-24, inversion 25, mirrors 40-42); a neighbour
 * outside the area is replaced by its image under the first operation, in
 * opsym's order, that takes it inside (the mirror groups: under the mirrors
 * of its negative coordinates), with that operation's character.  irrep:
 * -1 = the whole cluster, else a character vector: the operation negating
 * the coordinates of a set of axes has the character (-1)^(number of those
 * axes whose bit is set), and irrep is the smallest mask with the
 * representation's characters (for the mirror groups: the mirrors it is odd
 * under, a subset of the group's; -1: 0, 1; 2: 0, 1; 2/m: 0, 1, 4, 5;
 * 222: 0..3).  rvol = the point's number of images under the group / the
 * largest such number (the group's order once a grid point lies off every
 * plane and axis); the potential must have the group's symmetry.  Returns -1
 * for any other group or irrep. */
int fdmgpu_cluster_create_group(double rsort, double adimp, int lmaxso, int group, int irrep, fdmgpu_cluster **out);
/* The same with the choice of complex harmonics in the outer sphere
 * (the reference's Ylm_comp; mat.f90:1329, :1374, :1430-1436, :1476-1497):
 * the outer unknowns and incoming waves are Yc(L, m) (mat.f90:1733, the
 * exp(i m phi) combinations of the real harmonics) and the arithmetic is
 * complex.  Allowed for the groups whose operations keep every Yc(L, m)
 * (1 to 5; tab_data.f90:547) and required for the four-fold groups 4 (9),
 * -4 (10) and 4/m (11) (Atom_comp_cal, general.f90:9337), whose
 * representations are one-dimensional with complex characters (the
 * reference's Kari, calcMatRow's Repres_comp branch, mat.f90:1217-1223):
 * irrep = k + 4 [p = -1], the rotation by q pi/2 about z has the character
 * exp(-i k q pi/2) and a reflection of z adds the factor p (p = 1 for -4);
 * k = 0..3 for 4 and -4, 0..7 for 4/m.  Operations (opsym numbers): 4 =
 * {1, 18, 21, 24}, -4 = {1, 24, 28, 31}, 4/m = {1, 18, 21, 24, 25, 28, 31,
 * 42}; calculation areas and the mapping of neighbours as in posequiv
 * (general.f90:8980-8985).  The reference solves one representation of each
 * complex-conjugate pair and obtains the other's Taull by conjugation
 * (mat.f90:405-425); these systems are all built here.  Such clusters have
 * no device assembly tables (fdmgpu_cluster_assembly_tables returns -2):
 * assemble them on the host (fdmgpu_cluster_assemble, complex buffers). */
int fdmgpu_cluster_create_group_ylm(double rsort, double adimp, int lmaxso, int group, int irrep, int ylm_comp,
                                    fdmgpu_cluster **out);
/* A P1 cluster with natome muffin-tin atoms: posi[3*natome] (bohr, the
 * absorber at the origin, in non-decreasing distance from it, as the
 * reference sorts its cluster), sphere radii rmt[natome] (bohr), expansion
 * lmax[natome] (nlmsa = (lmax+1)^2 unknowns each, lmax <= 12).  It replaces
 * the atom branches of the set-up chain:
 *   reseau   general.f90:9563-9577   a point within rmt + epspos of an atom
 *                                    belongs to it (numia; the last atom in
 *                                    posi order wins) and leaves the grid
 *   bordure  general.f90:9924-10068  per atom, the nbordf interstitial points
 *                                    with a first-shell neighbour inside it
 *                                    (weights poidsa = 4 pi / nbordf), then the
 *                                    atom's points with an interstitial neighbour
 *   Ylmpt    general.f90:11533-11568 real harmonics about the atom at them
 *   newind   mat.f90:1790-1884       each atom's nlmsa rows go before the
 *                                    first interstitial point at least as far
 *                                    from the absorber as the atom (two atoms
 *                                    at one place follow each other), and the
 *                                    profile of the grid, atom and outer rows
 * Rows: the npr interstitial points in radial order with the atom blocks
 * inserted, then the nlmso outer rows (nligne = npr + sum nlmsa + nlmso).
 * Assemble with fdmgpu_cluster_assemble_atoms; fdmgpu_cluster_assemble and
 * the device assembler's tables return -2 for such a cluster.  Returns -1 on
 * bad arguments, unsorted atoms or an atom that touches no interstitial
 * point. */
int fdmgpu_cluster_create_atoms(double rsort, double adimp, int lmaxso, int natome, const double *posi,
                                const double *rmt, const int *lmax, fdmgpu_cluster **out);
/* The same cluster's rows for the irregular cluster solution (irregular = 1;
 * 0 gives fdmgpu_cluster_create_atoms), as mat (mat.f90:135-217) builds them
 * for FDM_comp with Test_reg = 0: newind (:1761-1921) with n_ato_RI = 2, so
 * each atom holds two blocks of nlmsa rows (the regular then the irregular
 * atomic solution's amplitudes, rows ianew .. ianew + 2 nlmsa - 1), and no
 * outer-sphere rows (nligne = npr + 2 sum nlmsa); a grid row with an
 * outside neighbour reaches to nligne in the profile, as newind sets it.
 * Assemble with fdmgpu_cluster_assemble_atoms_ri; the device assembler's
 * tables return -2 for it. */
int fdmgpu_cluster_create_atoms_ri(double rsort, double adimp, int lmaxso, int natome, const double *posi,
                                   const double *rmt, const int *lmax, int irregular, fdmgpu_cluster **out);
/* sizes[0..3] = natome, npr (interstitial points), sum of nbord, and the
 * length of phiato (sum of nbord * nlmsa) */
int fdmgpu_cluster_atom_sizes(const fdmgpu_cluster *c, int64_t *sizes);
/* Atom a (0-based): info[4] = nbord, nbordf, nlmsa, ianew (rows ianew ..
 * ianew + nlmsa - 1, 0-based, are its expansion); ibord[nbord] its border
 * points (0-based), the nbordf outside ones first; poidsa[nbordf];
 * ylmato[nbord * nlmsa] row-major.  Any output may be NULL. */
int fdmgpu_cluster_atom(const fdmgpu_cluster *c, int a, int *info, int *ibord, double *poidsa, double *ylmato);
/* numia[npoint]: 0, or 1 + the atom a point lies in; row[npoint]: the
 * 0-based row of an interstitial point, -1 for a point inside an atom.
 * Either may be NULL. */
int fdmgpu_cluster_point_rows(const fdmgpu_cluster *c, int *numia, int *row);
/* rvol of the npoint interior points (1 for P1) */
int fdmgpu_cluster_rvol(const fdmgpu_cluster *c, double *rvol);
/* harmonic index l*l + l + m of each of the nlmso outer unknowns */
int fdmgpu_cluster_harmonics(const fdmgpu_cluster *c, int *lm);
void fdmgpu_cluster_destroy(fdmgpu_cluster *c);
/* sizes[0..5] = npoint, npso, nligne, nlmso, nsortf, nsort */
int fdmgpu_cluster_sizes(const fdmgpu_cluster *c, int *sizes);
/* (x, y, z) of the npoint interior points in row order, bohr; xyz[3*npoint] */
int fdmgpu_cluster_points(const fdmgpu_cluster *c, double *xyz);
/* lb1/lb2[nligne], 1-based inclusive */
int fdmgpu_cluster_profile(const fdmgpu_cluster *c, int *lb1, int *lb2);
/* number of stored entries of the compressed rows */
int64_t fdmgpu_cluster_nnz(const fdmgpu_cluster *c);
/* One energy point: compressed rows (row_ptr[nligne+1] 0-based offsets,
 * col 1-based, values re/im) and the incoming-wave RHS calcMatRow adds
 * (rhs_re/rhs_im (nligne, nlmso) row-major; may be NULL).  V[npoint] in Ry.
 * im/rhs_im are only written when the arithmetic is complex. */
int fdmgpu_cluster_assemble(const fdmgpu_cluster *c, const double *V, double Enervide, double Eimag,
                            double Ecinetic_out, int E_comp, int basereel, int64_t *row_ptr, int *col,
                            double *re, double *im, double *rhs_re, double *rhs_im);
/* The same for a cluster with atoms (calcMatRow, mat.f90:1034-1632), given
 * the atoms' radial solutions at their border points, phiato: for each atom
 * in turn an (nbord, nlmsa) row-major block, phi_re and, for complex radial
 * solutions (the reference's nphiato7 = 2), phi_im (NULL: real; must be NULL
 * when the arithmetic is real).  Grid rows with a neighbour inside atom a
 * couple to a's expansion with -clapl(v) phiato at that neighbour
 * (:1233-1309); atom row Lm holds -poidsa Ylm in the columns of the nbordf
 * outside border points and sum_ib poidsa Ylm phiato(ib, Lm') in its own
 * block (:1515-1632); atom rows carry no incoming wave.  V[npoint] (the
 * values at points inside atoms are not used). */
int fdmgpu_cluster_assemble_atoms(const fdmgpu_cluster *c, const double *V, double Enervide, double Eimag,
                                  double Ecinetic_out, int E_comp, int basereel, const double *phi_re,
                                  const double *phi_im, int64_t *row_ptr, int *col, double *re, double *im,
                                  double *rhs_re, double *rhs_im);
/* The same with the irregular atomic solutions phi2 (same layout as phi),
 * for a cluster made with irregular = 1 (calcMatRow with irregular,
 * n_ato_RI = 2, Test_reg = 0): grid rows couple to both blocks of an atom
 * (:1266-1307); an outside neighbour adds no column and -(-i j_l + n_l) Ylm
 * times the stencil weight to the right-hand side (:1318-1398); atom rows of
 * the first block match the grid on the first nbordf/2 border points, those
 * of the second on the rest, and both hold sum_ib poidsa Ylm phi over the
 * first half in the regular block's columns and phi2 over the second half
 * in the irregular block's (:1515-1630).  Complex energy only (E_comp = 1,
 * basereel = 0).  A regular cluster takes phi2 = NULL (then it is
 * fdmgpu_cluster_assemble_atoms). */
int fdmgpu_cluster_assemble_atoms_ri(const fdmgpu_cluster *c, const double *V, double Enervide, double Eimag,
                                     double Ecinetic_out, int E_comp, int basereel, const double *phi_re,
                                     const double *phi_im, const double *phi2_re, const double *phi2_im,
                                     int64_t *row_ptr, int *col, double *re, double *im, double *rhs_re,
                                     double *rhs_im);
/* Scalar-relativistic rows (the reference's Relativiste = T; calcMatRow,
 * mat.f90:1155-1156, :1171-1185): fdmgpu_cluster_assemble and
 * fdmgpu_cluster_assemble_atoms_ri with the gradient of the potential at the
 * interior points, gradvr[npoint * 3] (point-major: dV/dx, dV/dy, dV/dz of
 * point i at 3 i .. 3 i + 2, Ry/bohr; the reference's gradvr(nicm, 3)).  A
 * grid row's diagonal loses a2s4 (V - E)^2 and the stencil coefficient of
 * neighbour v, before the character and rvol factors, becomes
 * -clapl(v) + bder cgrad(v) gradvr(i, k) with bder = -a2s4 / (1 - a2s4
 * (V - E)), a2s4 = alpha^2 / 4 (alpha = 0.0072973525664), cgrad laplac's
 * gradient stencil (general.f90:9791-9822) and k the axis of v; every use of
 * the coefficient (interior, outer-sphere and atom neighbours, the incoming
 * wave) takes the new value.  gradvr = NULL gives the plain rows. */
int fdmgpu_cluster_assemble_rel(const fdmgpu_cluster *c, const double *V, const double *gradvr, double Enervide,
                                double Eimag, double Ecinetic_out, int E_comp, int basereel, int64_t *row_ptr,
                                int *col, double *re, double *im, double *rhs_re, double *rhs_im);
int fdmgpu_cluster_assemble_atoms_rel(const fdmgpu_cluster *c, const double *V, const double *gradvr,
                                      double Enervide, double Eimag, double Ecinetic_out, int E_comp, int basereel,
                                      const double *phi_re, const double *phi_im, const double *phi2_re,
                                      const double *phi2_im, int64_t *row_ptr, int *col, double *re, double *im,
                                      double *rhs_re, double *rhs_im);
/* The device assembler's tables of the relativistic rows
 * (fdmgpu_rel_tables, fdmgpu.h): rq, rfac, rdiv [npoint*12]; opoint
 * [nrow_out], odir, ofac [nlink] (sizes of fdmgpu_cluster_assembly_sizes);
 * apoint [nar], adir [nal] (fdmgpu_cluster_atom_assembly_sizes; NULL without
 * atoms); clapl, cgrad [13].  Links in the order of the assembly and atom
 * tables.  Returns -2 for an irregular cluster. */
int fdmgpu_cluster_rel_tables(const fdmgpu_cluster *c, int64_t *rq, double *rfac, double *rdiv, int *opoint,
                              int *odir, double *ofac, int *apoint, int *adir, double *clapl, double *cgrad);

/* E-independent tables of one structure's per-energy assembly: the parts of
 * calcMatRow (mat.f90:1142-1232 grid rows, :1312-1398 outer columns,
 * :1405-1511 outer rows) that do not change with the energy, for a device
 * assembler (fdmgpu_assembler_create) to hold while only the energy-
 * dependent entries are recomputed per point.  Sizes:
 *   sizes[0] = nrow_out  grid rows with a stencil neighbour outside the sphere
 *   sizes[1] = nlink     (row, outside neighbour) couplings of those rows
 *   sizes[2] = nb        outer-sphere points carrying Bessel/Neumann functions
 *                        (the nsortf inner ones first, then the outside ones)
 *   sizes[3] = nnz */
int fdmgpu_cluster_assembly_sizes(const fdmgpu_cluster *c, int64_t *sizes);
/* base[nnz]      value of every stored entry with the energy terms left out:
 *                -clapl(v) for interior neighbours, -pw*Ylm for the outer
 *                rows' boundary columns, 0 where the value depends on E
 * qdiag[npoint]  position of grid row i's diagonal
 * orow[nrow_out] grid row (0-based) of each row with outside neighbours,
 * oq[nrow_out]   position of its first outer column (nlmso columns follow),
 * optr[nrow_out+1], ob[nlink], ow[nlink]: its outside neighbours in the
 *                reference's stencil order, as outer-point index and weight
 *                -clapl(v)
 * ylm[nb*nlmso]  real harmonics at the outer-sphere points
 * brad[nb]       their radii (bohr)
 * par[2]         clapl(0), outer-row quadrature weight pw = 4 pi / nsortf
 * For a representation of a mirror group the stencil weights carry the character and
 * rvol factors (-clapl(v) chi rvol(i), over rvol(j) in base) and the outer
 * rows' boundary columns are -poidso Ylm / rvol(j). */
int fdmgpu_cluster_assembly_tables(const fdmgpu_cluster *c, double *base, int64_t *qdiag, int *orow,
                                   int64_t *oq, int *optr, int *ob, double *ow, double *ylm, double *brad,
                                   double *par);
/* The same tables for a cluster with complex harmonics or characters
 * (Ylm_comp, the four-fold groups): the real parts where
 * fdmgpu_cluster_assembly_tables writes its values, the imaginary parts of
 * base (interior couplings chi -clapl(v) rvol(i) / rvol(j), the outer rows'
 * boundary columns -poidso conj(Yc) / rvol(j)), of the outside links'
 * weights ow and of the harmonics at the outer points (Yc) in base_im,
 * ow_im and ylm_im, and the terms of a point's own images (per axis, real
 * and imaginary, dself / dself_im [3*npoint]), which replace
 * fdmgpu_cluster_assembly_weights's dself.  -2 for a cluster with real
 * harmonics and characters, with atoms or irregular. */
int fdmgpu_cluster_assembly_tables_complex(const fdmgpu_cluster *c, double *base, double *base_im, int64_t *qdiag,
                                           int *orow, int64_t *oq, int *optr, int *ob, double *ow, double *ow_im,
                                           double *ylm, double *ylm_im, double *brad, double *par, double *dself,
                                           double *dself_im);
/* The rest of a device assembler's tables:
 * lharm[nlmso]    l of each outer unknown
 * wso[nsortf]     outer-row quadrature weight of each inner boundary point
 *                 (poidso, bordure general.f90:10073-10075; pw for P1)
 * dself[3*npoint] per axis, the stencil term of a point's own mirror image
 *                 (second-shell neighbour across a mirror plane), which adds
 *                 to its diagonal (0 for P1) */
int fdmgpu_cluster_assembly_weights(const fdmgpu_cluster *c, int *lharm, double *wso, double *dself);
/* For a cluster with atoms the three functions above describe rows, not
 * points: qdiag[i] = -1 for a point inside an atom, orow holds row numbers,
 * the outer rows are the last nlmso, and base also holds the atom rows'
 * -poidsa Ylmato in the columns of their outside border points
 * (mat.f90:1560-1581).  The entries that depend on phiato come from these
 * tables (fdmgpu_atom_tables in fdmgpu.h):
 * sizes[0..6] = natome, nar, nal, nphi, sum nlmsa (2 sum nlmsa for the
 * irregular solution's two row blocks), sum nbordf, sum nbordf * nlmsa.
 * For the irregular cluster solution (mat.f90:140-147) every atom has two
 * row and column blocks (regular then irregular atomic solutions, each row
 * block on its half of the outside border), bq lists both row blocks, the
 * grid rows touching the outside have no outer columns (oq = -1) and there
 * are no outer rows. */
int fdmgpu_cluster_atom_assembly_sizes(const fdmgpu_cluster *c, int64_t *sizes);
int fdmgpu_cluster_atom_assembly_tables(const fdmgpu_cluster *c, int64_t *aq, int *aa, int *aptr, int64_t *alphi,
                                        double *alw, int *ianew, int *nlmsa, int *nbordf, int64_t *phoff,
                                        int64_t *bq, double *poidsa, double *ylmato);

#ifdef __cplusplus
}
#endif
#endif
