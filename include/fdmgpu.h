/*
 * fdmgpu — MI355X (gfx950) solver for the FDMNES finite-difference system.
 *
 * C ABI of libfdmgpu.so.  Plain pointers and sizes only.  Every function
 * returns 0 on success, FDMGPU_PIVOT (1) when a vanishing pivot stops the
 * factorisation (the reference's STOP, see below), or a negative error code;
 * fdmgpu_last_error() describes the last failure of the calling thread.
 *
 * Conventions shared with the reference (mat_solve_gaussian.f90:5-65):
 *   - nligne unknowns (rows), nlmso right-hand sides;
 *   - lb1/lb2: 1-based inclusive column bounds of each row's profile
 *     (newind, mat.f90:1812-1884); lb1 must be non-decreasing;
 *   - smr/smi: RHS in, solution out, column-major smr(nlmso, nligne) —
 *     the RHS index is fastest — split into real and imaginary arrays;
 *     smi is ignored (may be NULL) when cal_comp == 0 (real arithmetic,
 *     mat.f90:96-104);
 *   - the matrix is passed as compressed rows gathered from calcMatRow's
 *     dense row buffers (mat.f90:1034): row_ptr[nligne+1] 0-based offsets,
 *     col[] 1-based column numbers, a_re[]/a_im[] values; every column
 *     must lie inside [lb1(ii), lb2(ii)].
 * Pivoting policy is the reference's: none.  Rows are eliminated from the
 * last to the first (a UL factorisation, A = U L with L unit lower); the
 * solve stops with FDMGPU_PIVOT and *fail_row = ii (1-based) at the first
 * row, in that order, whose pivot has |a_ii|^2 < 1e-20 (complex,
 * mat_solve_gaussian.f90:266-273) or |a_ii| < 1e-10 (real, :289-295).
 * On FDMGPU_PIVOT the RHS arrays are left unchanged.
 */
#ifndef FDMGPU_H
#define FDMGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FDMGPU_OK 0
#define FDMGPU_PIVOT 1
#define FDMGPU_EINVAL (-1)
#define FDMGPU_EHIP (-2)
#define FDMGPU_ENOMEM (-3)
#define FDMGPU_ESTATE (-4)

/* Precision of the factorisation. */
#define FDMGPU_FP64 0  /* complex128 / float64 factors */
#define FDMGPU_FP32 1  /* complex64 / float32 factors, FP64 refinement */

/* Library set-up on HIP device `device` (no reference counterpart: the
 * Fortran program has no device).  Idempotent. */
int fdmgpu_init(int device);
/* Releases cached device buffers and the drop-in's kept solvers. */
int fdmgpu_finalize(void);
const char *fdmgpu_last_error(void);
/* Name of the solver, 5 characters, as getSolverParams returns it
 * (mat_solve_gaussian.f90:376-389 returns 'Gauss'); writes "HIP  ". */
void fdmgpu_solver_name(char out[5]);

/*
 * Drop-in for the body of mat_solve (mat_solve_gaussian.f90:67-358): the
 * symbolic pass, the bottom-up elimination with the pivot test and the
 * forward substitution, in FP64, on the GPU.  Solves A x = b for all nlmso
 * right-hand sides; b is read from smr/smi and x written back in place.
 * The direct solve is followed by at most two correction sweeps with the
 * same FP64 factors while max_l ||r_l||_2/||b_l||_2 > 1e-12 (residual in
 * FP64 from the compressed rows): the blocked factorisation's explicit
 * 64 x 64 diagonal-tile inverses round worse than the reference's row by
 * row elimination on ill-conditioned real systems, and the corrections
 * bring the solution back to the reference's accuracy or better.
 * t_fill / t_tria (may be NULL) are incremented by the seconds spent in
 * device assembly and in factorisation + substitution, like the reference's
 * Time_fill / Time_tria accumulators (mat_solve_gaussian.f90:198-315).
 * The device solver of the last structure (profile and pattern compared
 * exactly) is kept for the next call; fdmgpu_finalize releases it.  Not
 * reentrant across threads (neither is mat_solve).
 */
int fdmgpu_mat_solve(int nligne, int nlmso, int cal_comp, const int *lb1, const int *lb2,
                     const int64_t *row_ptr, const int *col, const double *a_re, const double *a_im,
                     double *smr, double *smi, int *fail_row, double *t_fill, double *t_tria);

/*
 * Same contract as fdmgpu_mat_solve, factorising in FP32 (complex64 when
 * cal_comp) and recovering FP64 accuracy by iterative refinement: residual
 * r = b - A x and the update x += d in FP64 with the FP64 matrix, d from the
 * FP32 factors.  At most max_sweeps refinement sweeps; stops when
 * max over right-hand sides of ||r||_2/||b||_2 <= tol.  If refinement does
 * not reach tol (growth without pivoting can ruin FP32 factors), the system
 * is refactorised in FP64, refined with those factors in the same way, and
 * *sweeps is returned negative (-(FP32 sweeps done)).
 * No counterpart in the reference (it is FP64 only); same callers as
 * mat_solve.  *final_rel (may be NULL) receives the last residual ratio.
 */
int fdmgpu_mat_solve_mixed(int nligne, int nlmso, int cal_comp, const int *lb1, const int *lb2,
                           const int64_t *row_ptr, const int *col, const double *a_re, const double *a_im,
                           double *smr, double *smi, int max_sweeps, double tol, int *sweeps,
                           double *final_rel, int *fail_row);

/*
 * Device-resident solver for the energy loop (fdm.f90:3493-4192 calls mat
 * once per energy point; here many points are solved together).  A solver
 * holds up to 32 system slots of one precision and arithmetic; each slot is
 * a structure (profile lb1/lb2 + compressed-row pattern) with its own device
 * factor storage.  A solve call factorises and solves the listed slots
 * together — every kernel launch is grouped over them, so points of clusters
 * of uneven size share launches — and the launch sequence of each slot set
 * is captured once in a HIP graph and replayed.  Device pointers:
 *   d_vals[k]: matrix values of point k in the slot's pattern order,
 *              interleaved (re, im) when complex;
 *   d_b[k], d_x[k]: (nligne, nrhs) row-major, interleaved when complex
 *              (the smr/smi layout of mat_solve with re/im interleaved).
 * No counterpart in the reference (it solves one point at a time on the
 * CPU); callers: the energy loop (fdm.f90's loop over energy points,
 * taken a group of points at a time) and the batched benchmark.
 */
typedef struct fdmgpu_solver fdmgpu_solver;

/* precision FDMGPU_FP64 or FDMGPU_FP32; hip_stream may be NULL (library
 * stream) or a hipStream_t the solver launches on. */
int fdmgpu_solver_create(int precision, int cal_comp, int nrhs, void *hip_stream, fdmgpu_solver **out);
void fdmgpu_solver_destroy(fdmgpu_solver *s);
/* Adds a slot; allocates its factor storage. */
int fdmgpu_solver_add_system(fdmgpu_solver *s, int nligne, const int *lb1, const int *lb2, const int64_t *row_ptr,
                             const int *col, int *slot);
/* Solves count points on distinct slots.  FP32 solvers refine in FP64 (at
 * most max_sweeps sweeps to max_l ||r_l||/||b_l|| <= tol) and fall back to
 * FP64 factors per point when refinement does not converge (sweeps[k] < 0;
 * the FP64 solution is refined with its own factors to the same tol).
 * FP64 solvers with max_sweeps > 0 follow the direct solve by at most
 * max_sweeps correction sweeps with the same factors while
 * max_l ||r_l||/||b_l|| > tol (no fallback; sweeps[k] = corrections taken,
 * rel[k] = the final residual); max_sweeps <= 0: the direct solve only.
 * fail_rows[k] gets the reference's failing row (0 = none); returns
 * FDMGPU_PIVOT if any point failed.  sweeps/rel may be NULL. */
int fdmgpu_solver_solve(fdmgpu_solver *s, int count, const int *slots, const double *const *d_vals,
                        const double *const *d_b, double *const *d_x, int max_sweeps, double tol, int *fail_rows,
                        int *sweeps, double *rel);
/* Adjoint of sigma(E) = -Im sum_L s_L^T x_L over the diagonal potential:
 * with the factors left by the last solve of each slot (FP64 complex
 * solvers), solves A^T lam = s by the transposed substitutions and writes
 * d_grad[k][i] = Im sum_L lam_{i,L} x_{i,L} = d sigma / d V_i (nligne
 * entries).  No counterpart in the reference (it has no derivatives);
 * callers: potential fitting. */
int fdmgpu_solver_adjoint(fdmgpu_solver *s, int count, const int *slots, const double *const *d_src,
                          const double *const *d_x, double *const *d_grad);
/* sigma[k] = -Im sum_{i,L} src_{i,L} x_{i,L} (host output). */
int fdmgpu_solver_sigma(fdmgpu_solver *s, int count, const int *slots, const double *const *d_src,
                        const double *const *d_x, double *sigma);
/* The same sigma, bit for bit, written to device memory d_sigma[0..count-1]
 * with no allocation and no synchronisation: the kernels run on the solver's
 * stream (ordered after its solves) and keep their partial sums in a static
 * scratch array of the library (64 KiB per device), so count is at most 32
 * (FDMGPU_EINVAL beyond).  Every call on one device shares that scratch:
 * calls that overlap on one device (solvers on different streams, other
 * threads) race and must be ordered by the caller.  No counterpart in the
 * reference; callers: those that cannot allocate (work captured into a HIP
 * graph, tight serving loops). */
int fdmgpu_solver_sigma_noalloc(fdmgpu_solver *s, int count, const int *slots, const double *const *d_src,
                                const double *const *d_x, double *d_sigma);
/* Tuning: HIP-graph capture on/off, event timing of the trailing-update
 * launches on/off (disables graphs), workgroups of the persistent GEMM. */
int fdmgpu_solver_options(fdmgpu_solver *s, int use_graphs, int time_updates, int persist_wg);
/* Tile-schedule parameters (the autotuner's knobs): gk block steps per
 * trailing update (at most 32; effective panel width 64*gk, 0 = one-step updates),
 * workgroups of the persistent trailing-update grid, persistent (1) or
 * per-block (0) substitutions and their workgroup count, and the output
 * columns per workgroup of the group update (64, or 128: two tiles of a
 * block row); lookahead 1: at the end of a group, the update of the trailing
 * matrix outside the next group's block rows and columns runs on a second
 * stream beside the next group's panels (0: one update launch per group;
 * n > 1: as 1, that update on a persistent grid of n workgroups);
 * <= 0 keeps a value (< 0 for lookahead).  Takes effect at the next solve
 * (cached graphs are rebuilt). */
/* Overlap inside a trailing-update group (1 on, 0 off): the left-looking
 * update of block row and column K-1 from the group's steps K+1.. runs on a
 * second stream beside step K's panel and L-row/U-column products, and only
 * step K's own contribution follows them on the solver's stream.  No
 * counterpart in the reference (it has no such schedule).  Takes effect at
 * the next solve. */
int fdmgpu_solver_overlap(fdmgpu_solver *s, int on);
int fdmgpu_solver_tune(fdmgpu_solver *s, int gk, int trail_wg, int persist_solve, int solve_wg, int trail_tile,
                       int lookahead);
/* out[0] ms spent in trailing-update launches (when timed), out[1] their
 * count, out[2] multiply-adds of the trailing updates of one factorisation
 * of every slot at tile granularity, out[3] multiply-adds of the whole
 * factorisation on the element profile, out[4] factor bytes of all slots,
 * out[5] profile entries (solve multiply-adds per RHS), out[6] tile-level
 * multiply-adds of the panels and triangular block products (the part of
 * the factorisation outside the trailing updates), out[7] algorithmic bytes
 * of the timed trailing-update launches (every output tile read and
 * written once, every U and L tile of the step range read once). */
int fdmgpu_solver_stats(const fdmgpu_solver *s, double *out);
/* Test hooks, no counterpart in the reference: with snapshots on (1), each
 * solve keeps host copies of every solved slot's factor tiles and inverse
 * tiles after the factorisation and, per refinement sweep, of the FP32
 * correction and of the FP64 x after it (the last solve only; the copies
 * synchronise the stream, so a solve is slower).  fdmgpu_solver_snapshot
 * copies one of them to out (capacity bytes) and sets *bytes to its size
 * (out may be NULL to ask for the size): what 0 = factor then inverse tiles,
 * 1 = correction of sweep `sweep`, 2 = x after sweep `sweep`, 3 = *bytes
 * set to the number of sweeps kept.  Callers: determinism tests that
 * compare graph replays stage by stage. */
int fdmgpu_solver_snapshots(fdmgpu_solver *s, int on);
int fdmgpu_solver_snapshot(const fdmgpu_solver *s, int slot, int what, int sweep, void *out, int64_t capacity,
                           int64_t *bytes);

/*
 * Per-energy row assembly on the device, for the energy loop over one
 * structure.  Replaces, for a cluster without atoms, what calcMatRow
 * (mat.f90:1034-1727) computes row by row on the CPU for every energy:
 * the grid diagonal V - E - clapl(0) - i Eimag (mat.f90:1147-1150), the
 * outer columns and incoming-wave RHS of grid rows touching the outside
 * (:1312-1398), and the outer rows (:1401-1511), from the outer-sphere
 * functions of phiso (mat.f90:739-1028, recurrence sphere.f90:7226-7248).
 * The energy-independent tables (fdmgpu_cluster_assembly_tables in
 * fdmgpu_cluster.h fills them) are copied to the device once; each call
 * then writes the value arrays (and RHS) of up to 16 energy points.
 */
typedef struct fdmgpu_assembler fdmgpu_assembler;
/* Host pointers; sizes as fdmgpu_cluster_assembly_sizes reports them. */
typedef struct {
  int npoint, nligne, nlmso, lmaxso, nsortf, nb, nrow_out, nlink;
  int64_t nnz;
  double clapl0, pw;        /* stencil centre weight, outer-row quadrature weight */
  const int64_t *row_ptr;   /* [nligne+1] */
  const double *base;       /* [nnz] energy-independent part of every entry */
  const int64_t *qdiag;     /* [npoint] position of grid point i's diagonal (-1: inside an atom) */
  const int *orow;          /* [nrow_out] grid rows with outside neighbours (0-based row numbers) */
  const int64_t *oq;        /* [nrow_out] position of each one's first outer column */
  const int *optr;          /* [nrow_out+1] offsets of its neighbours in ob/ow */
  const int *ob;            /* [nlink] outer point of each outside neighbour */
  const double *ow;         /* [nlink] its stencil weight -clapl(v) */
  const double *ylm;        /* [nb*nlmso] real harmonics at the outer points */
  const double *brad;       /* [nb] radii of the outer points (bohr) */
  /* symmetry-reduced systems (fdmgpu_cluster_assembly_weights); NULL = P1:
   * l of l*l + l + m, pw for every point, no self terms */
  const int *lharm;         /* [nlmso] l of each outer unknown */
  const double *wso;        /* [nsortf] outer-row quadrature weights */
  const double *dself;      /* [3*npoint] mirror-image terms added to the diagonals */
} fdmgpu_assembly_tables;
/* max_points (1..16): energy points per call; hip_stream may be NULL (library
 * stream).  The tables are checked for consistency and copied. */
int fdmgpu_assembler_create(const fdmgpu_assembly_tables *t, int max_points, void *hip_stream,
                            fdmgpu_assembler **out);
void fdmgpu_assembler_destroy(fdmgpu_assembler *a);
/* count points: host arrays Enervide/Eimag/Ecinetic_out[count] (Ry), device
 * d_V[k] (npoint, Ry), outputs d_vals[k] (nnz values in row_ptr order,
 * interleaved (re, im) when complex) and d_rhs[k] ((nligne, nlmso) row-major,
 * interleaved when complex; NULL array or entry = not written).  Complex
 * arithmetic when E_comp or !basereel (mat.f90:96-104).  update_only != 0:
 * the buffers already hold the output of an earlier call of this assembler
 * with the same E_comp/basereel (any energy, any V); only the entries that
 * depend on the energy or V are rewritten. */
int fdmgpu_assembler_run(fdmgpu_assembler *a, int count, const double *Enervide, const double *Eimag,
                         const double *Ecinetic_out, int E_comp, int basereel, int update_only,
                         const double *const *d_V, double *const *d_vals, double *const *d_rhs);

/*
 * The same for a cluster with muffin-tin atoms (point group P1), whose
 * entries also depend on the atoms' radial solutions at their border points
 * (phiato, computed per energy by the caller's radial solver): the grid rows'
 * couplings to the expansion of an atom a stencil neighbour lies in,
 * -clapl(v) phiato(ib, Lm) summed over such neighbours in stencil order
 * (calcMatRow, mat.f90:1233-1309), and each atom row's own block
 * sum_ib poidsa(ib) Ylmato(ib, Lm) phiato(ib, Lm') over the nbordf outside
 * border points (:1583-1632).  With atoms the assembly tables' rows are the
 * cluster's rows (fdmgpu_cluster_assembly_tables fills them so): qdiag is -1
 * for a point inside an atom, orow holds row numbers, and the outer rows
 * are the last nlmso.  fdmgpu_cluster_atom_assembly_tables
 * (fdmgpu_cluster.h) fills this struct; sizes from
 * fdmgpu_cluster_atom_assembly_sizes.
 */
typedef struct {
  int natome;
  int nar;                  /* (grid row, atom) pairs: a stencil neighbour of the row lies in the atom */
  int nal;                  /* their (row, neighbour) links */
  int64_t nphi;             /* length of one energy point's phiato: sum over atoms of nbord * nlmsa */
  const int64_t *aq;        /* [nar] position of the atom block's first column in that grid row */
  const int *aa;            /* [nar] the atom (0-based) */
  const int *aptr;          /* [nar+1] offsets of its links in alphi/alw */
  const int64_t *alphi;     /* [nal] offset of the neighbour's nlmsa values in phiato */
  const double *alw;        /* [nal] stencil weight -clapl(v) */
  const int *ianew;         /* [natome] first row of the atom's expansion (0-based) */
  const int *nlmsa;         /* [natome] its unknowns (lm order, l*l + l + m) */
  const int *nbordf;        /* [natome] outside border points */
  const int64_t *phoff;     /* [natome] offset of its (nbord, nlmsa) row-major block in phiato */
  const int64_t *bq;        /* [sum nlmsa] per atom row, position of the atom block's first column */
  const double *poidsa;     /* [sum nbordf] border weights, atom after atom */
  const double *ylmato;     /* [sum nbordf*nlmsa] Ylm about the atom at them, (nbordf, nlmsa) per atom */
  int nri;                  /* 2: the irregular cluster solution (two blocks per atom), else 1 */
} fdmgpu_atom_tables;
/* As fdmgpu_assembler_create, with the atom tables (checked and copied).
 * nligne must be the interstitial rows + sum nlmsa + nlmso. */
int fdmgpu_assembler_create_atoms(const fdmgpu_assembly_tables *t, const fdmgpu_atom_tables *at, int max_points,
                                  void *hip_stream, fdmgpu_assembler **out);
/* As fdmgpu_assembler_run, with d_phi[k] the device phiato of point k
 * (nphi values; interleaved (re, im) when phi_cplx, the reference's
 * nphiato7 = 2, which needs complex arithmetic).  An assembler with atoms is
 * run only by this entry point (fdmgpu_assembler_run returns
 * FDMGPU_EINVAL for it). */
int fdmgpu_assembler_run_atoms(fdmgpu_assembler *a, int count, const double *Enervide, const double *Eimag,
                               const double *Ecinetic_out, int E_comp, int basereel, int update_only,
                               const double *const *d_V, const double *const *d_phi, int phi_cplx,
                               double *const *d_vals, double *const *d_rhs);

/*
 * Scalar-relativistic rows (the reference's Relativiste = T; calcMatRow,
 * mat.f90:1155-1156, :1171-1185): the grid diagonal loses a2s4 (V - E)^2 and
 * every stencil coefficient of grid point i becomes -clapl(v) + bder
 * cgrad(v) dV/dx_k(i), bder = -a2s4 / (1 - a2s4 (V - E)), a2s4 = alpha^2 / 4,
 * k the axis of neighbour v; the coefficient then carries the character
 * and rvol factors as before and enters the interior, outer-sphere and atom
 * couplings and the incoming wave.  These depend on V and E, so an
 * assembler given these tables recomputes them per call from the gradient
 * of the potential (fdmgpu_assembler_run_rel).  fdmgpu_cluster_rel_tables
 * (fdmgpu_cluster.h) fills the struct; its link order is that of the
 * assembly and atom tables.
 */
typedef struct {
  int npoint, nrow_out, nlink, nar, nal;
  double clapl[13];         /* clapl(0:12) */
  double cgrad[13];         /* cgrad(1:12) at [1..12], laplac's gradient stencil (general.f90:9791-9822) */
  const int64_t *rq;        /* [npoint*12] position written by neighbour v of point i (its column in i's row, the
                               diagonal for the point's own image), -1: outside the sphere or inside an atom */
  const double *rfac;       /* [npoint*12] character of the neighbour's operation times rvol(i) */
  const double *rdiv;       /* [npoint*12] rvol of the neighbour */
  const int *opoint;        /* [nrow_out] grid point of each row with outside neighbours */
  const int *odir;          /* [nlink] direction (0..11) of each outside link */
  const double *ofac;       /* [nlink] character times rvol(i) */
  const int *apoint;        /* [nar] grid point of each (row, atom) pair (NULL without atoms) */
  const int *adir;          /* [nal] direction of each atom link */
} fdmgpu_rel_tables;
/* Complex harmonics or characters (Ylm_comp, the four-fold groups 4, -4,
 * 4/m; calcMatRow mat.f90:1164-1231, :1328-1497): the assembler was created
 * from the real parts of fdmgpu_cluster_assembly_tables_complex; this adds
 * the imaginary parts base_im [nnz], ow_im [nlink], ylm_im [nb * nlmso] and
 * the points' own-image terms dself / dself_im [3 * npoint] (replacing the
 * real ones).  The assembler then runs in complex arithmetic only (not with
 * atoms or relativistic tables). */
int fdmgpu_assembler_set_complex(fdmgpu_assembler *a, const double *base_im, const double *ow_im,
                                 const double *ylm_im, const double *dself, const double *dself_im);
/* Adds the relativistic tables to an assembler (checked and copied); after
 * this the assembler is run only by fdmgpu_assembler_run_rel. */
int fdmgpu_assembler_set_relativistic(fdmgpu_assembler *a, const fdmgpu_rel_tables *t);
/* The irregular cluster solution (an assembler created with nri = 2):
 * d_phi2[k] the irregular atomic solutions' phiato (fdmgpu_sphere_run_ri);
 * complex energy and the complex basis only (E_comp, !basereel), as the
 * reference solves it (FDM_comp); the RHS is the incoming wave -i j + n of
 * mat.f90:1318-1398. */
int fdmgpu_assembler_run_ri(fdmgpu_assembler *a, int count, const double *Enervide, const double *Eimag,
                            const double *Ecinetic_out, int E_comp, int basereel, int update_only,
                            const double *const *d_V, const double *const *d_phi, const double *const *d_phi2,
                            int phi_cplx, double *const *d_vals, double *const *d_rhs);
/* As fdmgpu_assembler_run_atoms (d_phi NULL without atoms), with d_G[k] the
 * device gradient of point k's potential (npoint x 3, point-major, Ry/bohr). */
int fdmgpu_assembler_run_rel(fdmgpu_assembler *a, int count, const double *Enervide, const double *Eimag,
                             const double *Ecinetic_out, int E_comp, int basereel, int update_only,
                             const double *const *d_V, const double *const *d_G, const double *const *d_phi,
                             int phi_cplx, double *const *d_vals, double *const *d_rhs);

/*
 * The atoms' radial solutions at their border points, phiato (Sphere,
 * sphere.f90:8-210, per energy in the reference's energy loop,
 * fdm.f90:3733): per atom the radial Schrodinger equation on its radial
 * grid with a spherical potential (mod_V :216, coef_sch_rad :324,
 * Sch_radial :413: series at the origin, outward three-point recurrence),
 * matched to the normalised Bessel and Hankel functions at rmtg (Renormal
 * :931, cal_ampl :1157: amplitude Wronskout / W(u, j), Tau = -W(u, j) /
 * W(u, -i h)), then at each border point the solution of each L
 * interpolated quadratically at the point's distance (cal_phiato :3186,
 * f_interp2 tools.f90:504) times the real harmonic Ylmato.  One spin, no
 * spin-orbit or Hubbard term, the regular solution, Renorm = T; with
 * relativiste the scalar-relativistic coefficients of coef_sch_rad.
 * The output layout is the device assembler's phiato (per atom an
 * (nbord, nlmsa) row-major block, atoms concatenated), so a call can feed
 * fdmgpu_assembler_run_atoms directly.
 */
typedef struct fdmgpu_sphere fdmgpu_sphere;
typedef struct {
  int natome;
  const int *nr;            /* [natome] radial points of each atom (>= 4) */
  const double *r;          /* [sum nr] radial grids (bohr, increasing from r > 0), concatenated */
  const double *V;          /* [sum nr] spherical potentials (Ry) */
  const int *Z;             /* [natome] atomic numbers */
  const int *lmax;          /* [natome] (0..12); nlmsa = (lmax + 1)^2 */
  const double *rmtg;       /* [natome] matching radii (bohr) */
  const int *nbord;         /* [natome] border points */
  const double *bxyz;       /* [3 sum nbord] their positions (bohr), atom after atom */
  const double *pos;        /* [3 natome] atom positions (bohr) */
  const double *ylmato;     /* [sum nbord * nlmsa] real harmonics about the atom at them, (nbord, nlmsa) per atom */
  int relativiste;          /* the scalar-relativistic radial equation (Relativiste) */
  double V_intmax;          /* V is capped at V_intmax when V_intmax < 1000 (mod_V) */
} fdmgpu_sphere_tables;
/* max_points (1..16) energy points per call; hip_stream may be NULL. */
int fdmgpu_sphere_create(const fdmgpu_sphere_tables *t, int max_points, void *hip_stream, fdmgpu_sphere **out);
void fdmgpu_sphere_destroy(fdmgpu_sphere *s);
/* count points: host Ecinetic / Eimag / Enervide [count] (Ry); d_phi[k]:
 * device phiato of point k (sum nbord * nlmsa values; interleaved (re, im)
 * when complex_out, else the real parts, which are the solution when Eimag
 * = 0 and Ecinetic >= 1e-10); d_tau: NULL or per point a device array of
 * sum (lmax + 1) complex Tau (interleaved).  Asynchronous on the stream. */
int fdmgpu_sphere_run(fdmgpu_sphere *s, int count, const double *Ecinetic, const double *Eimag,
                      const double *Enervide, int complex_out, double *const *d_phi, double *const *d_tau);
/*
 * The absorbing atom's radial integrals of the golden rule (Radial_integral,
 * sphere.f90:1490, its call of radial_matrix :2209, as Tenseur_car makes it,
 * tensor.f90:185): with the same renormalised regular solution u_L as
 * phiato, rof(is, ip, L) = fac(ip) * integral_0^Rmtsd psii_is(r) r u_L(r) r^ip
 * dr by f_integr3 (tools.f90:308), fac = alfa_sf / 2, 1, Vecond / 2,
 * -Vecond^2 / 6 for ip = 0..3 with Vecond = alfa_sf / 2 * max(0.001, Energ +
 * Eseuil(is)); the imaginary part from Im u_L.  set_integrals names the
 * absorbing atom (0-based), its core functions psii [nbseuil][nr of the
 * atom] (the reference's psii(1:nr, is)), the edge energies Eseuil (Ry), the
 * integration radius Rmtsd (beyond the third radial point) and ip0..ip_max
 * (0..3).  run_integrals is fdmgpu_sphere_run with, per point, Energ and a
 * device array d_rof[k] of nbseuil * (ip_max - ip0 + 1) * (lmax + 1) complex
 * values (interleaved), index (is * nip + ip - ip0) * (lmax + 1) + L.
 */
/* fdmgpu_sphere_run with the irregular solution as well (Sch_radial_solsing,
 * sphere.f90:2591, as Sphere calls it with n_RI = 2, :152-162): -i r h_L(k r)
 * from the matching point outwards, integrated inwards by the same
 * three-point recurrence; d_phi2[k] its values at the border points in
 * d_phi's layout (the imaginary part zero for real radial solutions), the
 * phiato of the irregular cluster system (fdmgpu_cluster_assemble_atoms_ri's
 * phiato2).  Every atom needs a radial point beyond its matching point. */
int fdmgpu_sphere_run_ri(fdmgpu_sphere *s, int count, const double *Ecinetic, const double *Eimag,
                         const double *Enervide, int complex_out, double *const *d_phi, double *const *d_phi2,
                         double *const *d_tau);
int fdmgpu_sphere_set_integrals(fdmgpu_sphere *s, int atom, int nbseuil, const double *psii, const double *Eseuil,
                                double Rmtsd, int ip0, int ip_max);
int fdmgpu_sphere_run_integrals(fdmgpu_sphere *s, int count, const double *Ecinetic, const double *Eimag,
                                const double *Enervide, const double *Energ, int complex_out, double *const *d_phi,
                                double *const *d_tau, double *const *d_rof);

/*
 * The absorbing atom's Cartesian cross-section tensors per energy point
 * (Tenseur_car, tensor.f90:15-946, with tens_ab :947-1345; called by the
 * energy loop at fdm.f90:3975): the electric dipole-dipole secdd(3, 3),
 * dipole-quadrupole secdq(3, 3, 3) and quadrupole-quadrupole secqq(3, 3, 3,
 * 3) tensors of every resolved core state or edge (ninitr), from the
 * atom's Taull and radial integrals rof:
 *   Ten(lm_s, lm_i) = i * sum Cg * (conj(rof1) rof2 Taull(lm1, lm2)
 *                     - conj(conj(rof2) rof1 Taull(lm2, lm1))) / 2
 * over core states and final-state harmonics (Cg = Clebsch-Gordan weights
 * times Gaunt coefficients), secXX = sum clm clm Ten over the spherical
 * components of the Cartesian operators, spread to the elements equal by
 * symmetry (ldip, lqua and the msym tables, as the reference's Tensor_shape
 * and tensdd / tensdq / tensqq give them), then turned by rot_int x
 * rot_atom_abs^T.  The branch of an FDM absorption calculation: no TDDFT,
 * optic, NRIXS, Green's-function or FDM_comp output, no singular solution,
 * no spin-orbit, one final-state spin, spherical potential, real harmonics,
 * electric multipoles E1E1, E1E2, E2E2.
 */
typedef struct fdmgpu_tensor fdmgpu_tensor;
typedef struct {
  int jseuil, lseuil, nbseuil; /* the edge (coef_init's Jseuil, Lseuil) and the number of edges (1, 2) */
  int core_resolved;           /* one tensor per core state (else per edge) */
  int lmax;                    /* the probe's lmax (0..3), nlm_probe = (lmax + 1)^2 */
  int lplus1, lmoins1;         /* final states l > lseuil only / l <= lseuil only */
  int numat;                   /* atomic number (1 halves the tensors: one core state) */
  int e1e1, e1e2, e2e2;        /* the multipoles computed */
  const int *ldip;             /* [3] */
  const int *lqua;             /* [3][3] */
  const int *msymdd, *msymddi; /* [3][3] */
  const int *msymdq, *msymdqi; /* [3][3][3] */
  const int *msymqq, *msymqqi; /* [3][3][3][3]: tables in C order of the reference's Fortran indices */
  const double *rot_int, *rot_atom_abs; /* [3][3], (i, j) at [3 i + j] */
  int rof_lmax, rof_nbseuil, rof_ip0, rof_ip_max; /* the layout of fdmgpu_sphere_run_integrals's rof */
  int64_t taull_offset;        /* complex offset of the absorber's block in a point's tau */
  int taull_ld;                /* its row length (the absorber's nlmsa) */
} fdmgpu_tensor_tables;
/* max_points 1..32; hip_stream may be NULL.  ninitr: tensors per point. */
int fdmgpu_tensor_create(const fdmgpu_tensor_tables *t, int max_points, void *hip_stream, fdmgpu_tensor **out);
void fdmgpu_tensor_destroy(fdmgpu_tensor *h);
int fdmgpu_tensor_ninitr(const fdmgpu_tensor *h);
/* count points: d_rof[k] (fdmgpu_sphere_run_integrals's output), d_tau[k]
 * (fdmgpu_tau_reduce's, after fdmgpu_tau_symmetrize as mat does); outputs
 * per point, complex interleaved, C order: d_secdd[k] [ninitr][3][3],
 * d_secdq[k] [ninitr][3][3][3], d_secqq[k] [ninitr][3][3][3][3] (NULL for a
 * multipole not computed).  Asynchronous on the stream. */
int fdmgpu_tensor_run(fdmgpu_tensor *h, int count, const double *const *d_rof, const double *const *d_tau,
                      double *const *d_secdd, double *const *d_secdq, double *const *d_secqq);

/*
 * Post-solve reduction to the atom amplitude products (mat.f90:297-405,
 * regular solution, nspino = 1, one representation), on solutions already
 * on the device.  Replaces the CPU loop after mat_solve in mat():
 *   - real basis (basereel): norm(Lm2, Lm) = delta + i x(nligneso + Lm, Lm2)
 *     (mat.f90:314-328), inverted by invcomp's ZGETRF/ZGETRI sequence
 *     (mat.f90:1927-1987, is = 0; sub_util.f:11149, 11382);
 *   - amplitudes a(Lmf, ii) = sum_k norm^-1(Lmf, k) x(ii, k) (norm_t,
 *     mat.f90:340, :368-376), or x(ii, Lmf) in the complex basis;
 *   - tau_ia(Lm1, Lm2) (+)= -i sum_Lmf a(Lmf, ianew+Lm1) conj(a(Lmf,
 *     ianew+Lm2)), or -i a a without the conjugate when fdm_comp
 *     (mat.f90:392-399; Taull(Lm01, is1, Lm02, is2, ia) with Lm01/Lm02 the
 *     caller's map of the atom row order).
 */
typedef struct {
  int nligne, nlmso;
  int nligneso;      /* 0-based first outer-sphere row (nligne - nlmso for the regular solution) */
  int cal_comp;      /* solutions complex (interleaved) or real */
  int basereel;      /* renormalise by the inverted outer-sphere block */
  int fdm_comp;      /* no conjugate in the contraction */
  int natome;
  const int *ianew;  /* host [natome]: 0-based first row of each atom's block */
  const int *nlmsa;  /* host [natome]: its row count */
} fdmgpu_tau_desc;
/* count energy points (any number; launched 16 at a time): d_x[k] device
 * solution (nligne, nlmso) row-major; d_tau[k] device output, per atom
 * nlmsa^2 complex entries [Lm1][Lm2] (interleaved), atoms concatenated,
 * added to when accumulate != 0 (Taull accumulates over representations and
 * spins); d_norm NULL or d_norm[k] NULL, else the inverted norm (nlmso^2
 * complex, column-major, as invcomp leaves it); d_info NULL or a device
 * array of count ints: 0, or the 1-based zero pivot of ZGETF2 at which the
 * reference stops (that point's tau and norm are then not written).
 * Asynchronous on hip_stream (NULL = library stream).  Limits: nlmso <= 256
 * and (nlmso^2 + nlmso*(max nlmsa + 1)) * 16 + 4*nlmso bytes <= 160 KiB. */
int fdmgpu_tau_reduce(const fdmgpu_tau_desc *d, int count, const double *const *d_x, double *const *d_tau,
                      double *const *d_norm, int accumulate, int *d_info, void *hip_stream);

/*
 * Tau_RI with complex energy and FDM_comp (mat.f90:229-295, Test_reg = 0,
 * one spin, one representation), from the regular cluster solution and the
 * irregular one (the system of fdmgpu_cluster_create_atoms_ri), both complex
 * and already on the device.  Replaces mat's loop after its two mat_solve
 * calls: per atom, Tau_RI(Lm1, Lm2, i_RI) += sum_Lmf Ampl(Lmf, Lm1)
 * Ampl_I(Lmf, (i_RI-1) nlmsa + Lm2), each product added in Lmf order as
 * the reference adds it, with Ampl the regular solution's atom rows rounded
 * to single precision (the reference keeps the regular solution as
 * cmplx(smr, smi) of default kind, mat.f90:211) and Ampl_I the irregular
 * solution's atom rows (the first n_ato_ri * nlmsa of the atom's blocks).
 */
typedef struct {
  int nlmso;
  int natome;
  const int *ianew_reg;  /* host [natome]: 0-based first row of each atom in the regular solution */
  const int *ianew_irr;  /* host [natome]: the same in the irregular solution */
  const int *nlmsa;      /* host [natome] */
  int nligne_reg, nligne_irr;
  int n_ato_ri;          /* atom blocks of the irregular system per atom: 2 (Test_reg = 0) */
  int n_ri;              /* 2 */
} fdmgpu_tau_ri_desc;
/* d_xreg[k], d_xirr[k]: (nligne, nlmso) interleaved complex; d_tau_ri[k]:
 * per atom n_ri * nlmsa^2 complex entries [i_RI][Lm1][Lm2], atoms
 * concatenated, added to when accumulate != 0.  Asynchronous on hip_stream
 * (NULL = library stream). */
int fdmgpu_tau_ri(const fdmgpu_tau_ri_desc *d, int count, const double *const *d_xreg, const double *const *d_xirr,
                  double *const *d_tau_ri, int accumulate, void *hip_stream);
/* The symmetrisation that ends mat (mat.f90:596-641, one spin, no
 * spin-orbit, not with FDM_comp), in place on the per-atom blocks written
 * by fdmgpu_tau_reduce: harmonics in (L, m) order (nlmsa = (lmax+1)^2);
 * ylm_comp != 0: each pair and its partner (L2 -m2, L1 -m1) become
 * c = 0.5 (T + (-1)^(m1+m2) T_partner) and (-1)^(m1+m2) c, as the
 * reference's sweep leaves them; else i Im(T) (real harmonics). */
int fdmgpu_tau_symmetrize(int natome, const int *nlmsa, int ylm_comp, int count, double *const *d_tau,
                          void *hip_stream);

#ifdef __cplusplus
}
#endif
#endif
