"""Synthetic workloads: clusters, potentials, sources and energy grids.

Each workload is a spherical FDM cluster (cubic grid, 4th-order stencil,
outer-sphere expansion of (lmaxso+1)^2 harmonics, no atoms, point group P1)
carrying a model potential, with right-hand sides that are real Y_lm x
Gaussian sources at the absorber:

  V(r_i)  = sum_atoms -2 Z exp(-d/lambda) / max(d, h/2)  (Ry, d in bohr)
            + uniform noise in [-0.1, 0.1] Ry (numpy PCG64, per-workload seed)
            The h/2 floor regularises the well of an atom sitting on a grid
            point (the absorber at the origin always does).
  s_L(i)  = Y_L(r^_i) exp(-r_i^2 / (2 sigma^2)),  sigma = 0.3 Angstrom
            (Y_L real harmonics, lmax = lmaxso, value on the +z axis at r = 0)
  sigma(E) = -Im sum_L s_L^T x_L  with  A(E) x_L = s_L

The system matrix is the reference's calcMatRow matrix for such a cluster
(mat.f90:1142-1511).  Energies are Enervide in Ry measured from the vacuum
level; the kinetic energy outside is Ecinetic_out = max(E, Eclie_out) with
Eclie_out = 1 eV (fdm.f90:3567-3572, lecture.f90:2013).  A non-zero
broadening Eimag makes the arithmetic complex and selects the complex
outer basis (mat.f90:96-104).
"""
from __future__ import annotations

import dataclasses
import itertools

import numpy as np

BOHR = 0.52917721067
RYDB = 13.605693012905


@dataclasses.dataclass(frozen=True)
class Workload:
    name: str
    lattice: str          # "fcc" | "bcc"
    a: float              # lattice constant, Angstrom
    Z: int
    atoms_radius: float   # atoms within this distance of the absorber, Angstrom
    R: float              # cluster (FDM sphere) radius, Angstrom
    h: float              # grid step, Angstrom
    lmaxso: int           # outer expansion / source lmax -> (lmaxso+1)^2 RHS
    gamma_half_eV: float  # imaginary broadening Gamma/2, eV (0 = real arithmetic)
    e_min_eV: float
    e_max_eV: float
    n_energies: int
    seed: int = 0
    screening: float = 0.5  # Angstrom
    source_width: float = 0.3  # Angstrom
    noise_Ry: float = 0.1

    @property
    def energies_Ry(self) -> np.ndarray:
        return np.linspace(self.e_min_eV, self.e_max_eV, self.n_energies) / RYDB

    @property
    def eimag_Ry(self) -> float:
        return self.gamma_half_eV / RYDB

    @property
    def complex(self) -> bool:
        return self.gamma_half_eV != 0.0


# The five workloads (Cu K-edge CPU case; Fe drop-in; three-cluster batch;
# mixed precision and adjoint on the Fe cluster).
CU_REAL = Workload("cu13_R3.0_h0.24_real", "fcc", 3.61, 29, 2.56, 3.0, 0.24, 2, 0.0, -5.0, 40.0, 91, seed=0)
FE_FP64 = Workload("fe59_R5.0_h0.2_c128", "bcc", 2.866, 26, 5.0, 5.0, 0.2, 3, 0.5, -10.0, 60.0, 200, seed=0)
BATCH = [
    Workload("fe_R3.0_h0.2_c128", "bcc", 2.866, 26, 3.0, 3.0, 0.2, 3, 0.5, -10.0, 60.0, 86, seed=1),
    Workload("fe_R4.0_h0.2_c128", "bcc", 2.866, 26, 4.0, 4.0, 0.2, 3, 0.5, -10.0, 60.0, 85, seed=2),
    Workload("fe_R5.0_h0.2_c128", "bcc", 2.866, 26, 5.0, 5.0, 0.2, 3, 0.5, -10.0, 60.0, 85, seed=3),
]
FE_MIXED = FE_FP64
FE_ADJOINT = FE_FP64
WORKLOADS = {w.name: w for w in [CU_REAL, FE_FP64, *BATCH]}


def lattice_sites(kind: str, a: float, radius: float) -> np.ndarray:
    """Lattice sites (Angstrom) within ``radius`` of the origin, origin first."""
    basis = {"fcc": [(0, 0, 0), (0.5, 0.5, 0), (0.5, 0, 0.5), (0, 0.5, 0.5)],
             "bcc": [(0, 0, 0), (0.5, 0.5, 0.5)]}[kind]
    m = int(np.ceil(radius / a)) + 1
    pts = []
    for i, j, k in itertools.product(range(-m, m + 1), repeat=3):
        for b in basis:
            p = a * (np.array([i, j, k], float) + np.array(b))
            if np.linalg.norm(p) <= radius + 1e-6:
                pts.append(p)
    pts = np.array(pts)
    order = np.lexsort((pts[:, 2], pts[:, 1], pts[:, 0], np.round(np.linalg.norm(pts, axis=1), 9)))
    return pts[order]


def potential(w: Workload, points_bohr: np.ndarray, shifts: np.ndarray | None = None,
              mirror_symmetric: bool = False) -> np.ndarray:
    """V on the interior points (Ry).  ``shifts`` (per atom, Ry) adds
    shift_a * exp(-d/lambda) around atom a (the per-atom parameters of the
    adjoint workload).  ``mirror_symmetric``: the noise term is drawn per
    (|ix|, |iy|, |iz|), so V keeps the lattice's mirror planes x, y, z = 0
    (point group mmm) and its values on the wedge x, y, z >= 0 define it."""
    sites = lattice_sites(w.lattice, w.a, w.atoms_radius) / BOHR
    lam = w.screening / BOHR
    dmin = 0.5 * w.h / BOHR
    V = np.zeros(len(points_bohr))
    for a, s in enumerate(sites):
        d = np.linalg.norm(points_bohr - s, axis=1)
        e = np.exp(-d / lam)
        V += -2.0 * w.Z * e / np.maximum(d, dmin)
        if shifts is not None:
            V += shifts[a] * e
    rng = np.random.Generator(np.random.PCG64(w.seed))
    if mirror_symmetric:
        idx = np.abs(np.rint(points_bohr / (w.h / BOHR)).astype(np.int64))
        n = int(np.ceil(w.R / w.h)) + 3
        V += rng.uniform(-w.noise_Ry, w.noise_Ry, (n, n, n))[idx[:, 0], idx[:, 1], idx[:, 2]]
    else:
        V += rng.uniform(-w.noise_Ry, w.noise_Ry, len(points_bohr))
    return V


def atom_weights(w: Workload, points_bohr: np.ndarray) -> np.ndarray:
    """d V_i / d shift_a = exp(-d_ia / lambda); shape (natoms, npoint)."""
    sites = lattice_sites(w.lattice, w.a, w.atoms_radius) / BOHR
    lam = w.screening / BOHR
    return np.stack([np.exp(-np.linalg.norm(points_bohr - s, axis=1) / lam) for s in sites])


def real_ylm(lmax: int, v: np.ndarray) -> np.ndarray:
    """Real spherical harmonics (tesseral, no Condon-Shortley phase) of the
    directions v (npts, 3); columns l*l + l + m.  On the z axis, and at the
    origin, the direction is taken as +z."""
    x, y, z = v[:, 0], v[:, 1], v[:, 2]
    r = np.sqrt(x * x + y * y + z * z)
    rho = np.sqrt(x * x + y * y)
    ct = np.where(r > 1e-6, z / np.where(r > 0, r, 1), 1.0)
    st = np.where(rho > 1e-6, np.sqrt(np.maximum(1 - ct * ct, 0)), 0.0)
    phi = np.arctan2(y, x)
    nl = lmax + 1
    P = np.zeros((nl, nl, len(v)))
    P[0, 0] = 1 / np.sqrt(4 * np.pi)
    for m in range(1, nl):
        P[m, m] = np.sqrt((2 * m + 1) / (2 * m)) * st * P[m - 1, m - 1]
    for m in range(0, lmax):
        P[m + 1, m] = np.sqrt(2 * m + 3) * ct * P[m, m]
    for m in range(nl):
        for l in range(m + 2, nl):
            a = np.sqrt((4 * l * l - 1) / (l * l - m * m))
            b = np.sqrt(((l - 1) ** 2 - m * m) / (4 * (l - 1) ** 2 - 1))
            P[l, m] = a * (ct * P[l - 1, m] - b * P[l - 2, m])
    out = np.zeros((len(v), nl * nl))
    for l in range(nl):
        out[:, l * l + l] = P[l, 0]
        for m in range(1, l + 1):
            out[:, l * l + l + m] = np.sqrt(2) * P[l, m] * np.cos(m * phi) * (st > 0)
            out[:, l * l + l - m] = np.sqrt(2) * P[l, m] * np.sin(m * phi) * (st > 0)
    return out


def sources(w: Workload, points_bohr: np.ndarray, nligne: int) -> np.ndarray:
    """s (nligne, (lmaxso+1)^2): Y_L x Gaussian on the grid rows, zero on
    the outer rows."""
    pa = points_bohr * BOHR
    r2 = np.sum(pa * pa, axis=1)
    g = np.exp(-r2 / (2 * w.source_width ** 2))
    s = np.zeros((nligne, (w.lmaxso + 1) ** 2))
    s[: len(pa)] = real_ylm(w.lmaxso, pa) * g[:, None]
    return s


def ecinetic_out(E: float) -> float:
    return max(E, 1.0 / RYDB)


def sigma(s: np.ndarray, x: np.ndarray) -> float:
    return float(-np.imag(np.sum(s * x)))
