"""Writes the golden vectors of tests/golden/ from the reference itself.

Runs the reference FDMNES routines compiled from their own Fortran sources
(oracle/_ref, built by ``make -f oracle/Makefile.ref``) on small synthetic
clusters and stores inputs and outputs as .npz fixtures:
  newind (mat.f90:1761)  -> lb1, lb2, newinv, new
  phiso (mat.f90:739)    -> Bessel / Neuman at the outer-sphere points
  mat_solve (mat_solve_gaussian.f90:5, with the reference calcMatRow)
                          -> solution smr/smi, for the three branch sets
                             (complex energy; real; complex arithmetic at real E)
  cbessneu / cbessneur (sphere.f90:7226/7356), cylm + ylmcr (general.f90:11606/11692)
  reseau / laplac / bordure (general.f90:9439/9732/9895, lists read back by
  recup_bordure)          -> point order, neighbour table, stencil weights,
                             outer-sphere boundary points and weights, for
                             point group 1, the subgroups of mmm (m, mm2,
                             mmm, -1, 2, 2/m, 222) and 4, -4, 4/m
                             (calculation area, rvol, symmetry operation of
                             each neighbour)
  the same with muffin-tin atoms, then newind and mat_solve with a synthetic
  phiato (j_L(k r) Ylm at the atoms' border points)
                          -> atom border lists, weights, Ylmato, atom blocks
                             in the profile, and the solution for the three
                             branch sets
  mat_solve with scalar-relativistic rows (Relativiste, gradvr the
  potential's gradient, laplac's cgrad), without and with the atoms
                          -> the solution for the three branch sets
  mat_solve with complex harmonics in the outer sphere (Ylm_comp)
                          -> the solution for three branch sets
  Sphere (sphere.f90:8) for each atom of that cluster
                          -> phiato at the border points and Tau, complex
                             and real energy, with and without Relativiste
  mat (mat.f90:5) on that cluster
                          -> Taull after its symmetrisation (real and complex
                             basis), and with FDM_comp at complex energy the
                             irregular cluster solution (newind and mat_solve
                             with irregular, n_ato_RI = 2) and Tau_RI
The cluster arrays fed to the reference (grid, neighbours, boundary lists,
Ylmso) come from the oracle's restatement; the fixtures store the profile
and matrix-relevant inputs so the tests can rebuild each case.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402
from oracle import ref as R  # noqa: E402

BOHR = 0.52917721067

CASES = [
    # name, R (A), h (A), lmaxso, E (Ry), Eimag (Ry), E_comp, cal_comp, basereel, seed
    ("c128_R1.2_h0.25", 1.2, 0.25, 2, 0.7, 0.05, True, True, False, 0),
    ("real_R1.2_h0.25", 1.2, 0.25, 2, 0.7, 0.0, False, False, True, 1),
    ("cal_R1.2_h0.25", 1.2, 0.25, 2, 0.7, 0.0, False, True, False, 2),
    ("c128_R1.6_h0.3_l3", 1.6, 0.3, 3, 1.3, 0.0367, True, True, False, 3),
]


GRID_CASES = [(1.2, 0.25), (1.6, 0.3), (2.0, 0.25)]
# point group -> its operations in the reference's numbering (opsym,
# general.f90:4587): identity 1, four-fold rotations about z 18 and 21,
# two-fold rotations about x, y, z 22-24, inversion 25, four-fold
# rotation-reflections about z 28 and 31, mirrors normal to x, y, z 40-42
GROUP_OPS = {3: [1, 42], 6: [1, 24, 40, 41], 8: [1, 22, 23, 24, 25, 40, 41, 42],
             2: [1, 25], 4: [1, 24], 5: [1, 24, 25, 42], 7: [1, 22, 23, 24],
             9: [1, 18, 21, 24], 10: [1, 24, 28, 31], 11: [1, 18, 21, 24, 25, 28, 31, 42]}


def grid_fixtures():
    for Ra, h in GRID_CASES:
        g = R.grid(Ra / BOHR, h / BOHR)
        np.savez_compressed(os.path.join(HERE, f"grid_p1_R{Ra}_h{h}.npz"), R=Ra, h=h, npoint=g["npoint"],
                            npso=g["npso"], indice=g["indice"], radius=g["radius"], ivois=g["ivois"],
                            clapl=g["clapl"], nsortf=g["nsortf"], nsort=g["nsort"], isrt=g["isrt"],
                            poidso=g["poidso"])
        print("grid P1", Ra, h, g["npoint"], g["npso"], g["nsortf"], g["nsort"])
    for grp, ops in GROUP_OPS.items():
        io = np.zeros(R.NOPSM, np.int32)
        io[np.array(ops) - 1] = 1
        Ra, h = 1.6, 0.25
        g = R.grid(Ra / BOHR, h / BOHR, igrpt=grp, iopsymr=io)
        npt = g["npoint"]
        opmat = np.array([R.opsym(op) for op in ops])
        np.savez_compressed(os.path.join(HERE, f"grid_g{grp}_R{Ra}_h{h}.npz"), R=Ra, h=h, group=grp, ops=np.array(ops),
                            opmat=opmat,
                            npoint=npt, npso=g["npso"], indice=g["indice"], radius=g["radius"], rvol=g["rvol"],
                            ivois=g["ivois"][:npt], isvois=g["isvois"][:npt], clapl=g["clapl"], nsortf=g["nsortf"],
                            nsort=g["nsort"], isrt=g["isrt"], poidso=g["poidso"])
        print("grid group", grp, npt, g["npso"], g["nsortf"], g["nsort"])


# a cluster with three atoms: the absorber and two at the same distance
# (newind's shared insertion point), positions in Angstrom
ATOMS = dict(R=1.9, h=0.25, lmaxso=2, posi=[[0.0, 0.0, 0.0], [1.0, 0.25, 0.0], [-1.0, -0.25, 0.0]],
             rmt=[0.45, 0.4, 0.4], lmax=[2, 1, 1])
ATOM_BRANCHES = [  # E (Ry), Eimag, E_comp, cal_comp, basereel
    (0.7, 0.05, True, True, False), (0.7, 0.0, False, False, True), (0.7, 0.0, False, True, False)]
ROW_STRIDE = 3


def atom_phiato(s, E, Eimag, E_comp):
    """Synthetic phiato: j_L(k r) Y_Lm at every border point of every atom,
    k = sqrt(E + 1.5 (+ i Eimag)), as the radial solutions of a flat well."""
    from scipy.special import spherical_jn
    k = np.sqrt(E + 1.5 + (1j * Eimag if E_comp else 0))
    out = []
    for ia in range(s["natome"]):
        nb, nl = int(s["nbord"][ia]), int(s["nlmsa"][ia])
        pts = s["xyz"][:3, s["ibord"][:nb, ia] - 1].T - s["posi"][:, ia]
        r = np.linalg.norm(pts, axis=1)
        ls = np.array([L for L in range(int(s["lmax"][ia]) + 1) for m in range(-L, L + 1)])
        phi = np.stack([spherical_jn(L, k * r) for L in ls], axis=1) * s["ylmato"][:nb, :nl, ia]
        out.append(phi if E_comp else phi.real)
    return out


def atom_fixture():
    a = ATOMS
    posi = np.array(a["posi"]) / BOHR
    s = R.atom_system(a["R"] / BOHR, a["h"] / BOHR, a["lmaxso"], posi, np.array(a["rmt"]) / BOHR, a["lmax"])
    rng = np.random.default_rng(7)
    V = rng.uniform(-0.3, 0.3, s["npoint"]) - 0.8
    rows = np.arange(0, s["nligne"], ROW_STRIDE)
    out = dict(R=a["R"], h=a["h"], lmaxso=a["lmaxso"], posi=np.array(a["posi"]), rmt=np.array(a["rmt"]),
               lmax=np.array(a["lmax"]), V=V, rows=rows, npoint=s["npoint"], npr=s["npr"], nligne=s["nligne"],
               numia=s["numia"][: s["npso"]], ianew=s["ianew"], nbord=s["nbord"], nbordf=s["nbordf"],
               lb1=s["lb1"], lb2=s["lb2"], newinv=s["newinv"], new=s["new"])
    for ia in range(s["natome"]):
        nb, nl = int(s["nbord"][ia]), int(s["nlmsa"][ia])
        out[f"ibord{ia}"] = s["ibord"][:nb, ia]
        out[f"poidsa{ia}"] = s["poidsa"][: s["nbordf"][ia], ia]
        out[f"ylmato{ia}"] = s["ylmato"][:nb, :nl, ia]
    for b, (E, Ei, E_comp, cal, basereel) in enumerate(ATOM_BRANCHES):
        bes, neu = R.phiso_sys(s, E_comp, E, Ei, E)
        x = R.mat_solve_sys(s, V, E, Ei, bes, neu, E_comp, cal, basereel, atom_phiato(s, E, Ei, E_comp))
        out[f"x{b}"] = x[rows]
    np.savez_compressed(os.path.join(HERE, "atoms_R1.9_h0.25.npz"), **out)
    print("atoms", s["npoint"], s["npr"], s["nligne"], s["nbord"], s["nbordf"], s["ianew"])


def atom_phiato_irregular(s, E, Eimag):
    """Synthetic irregular atomic solutions: y_L(k r) Y_Lm at every border
    point, k = sqrt(E + 1.5 + i Eimag), r clamped to rmt / 2 (the atom's
    centre point can be a border point, where y_L diverges)."""
    from scipy.special import spherical_yn
    k = np.sqrt(E + 1.5 + 1j * Eimag)
    out = []
    for ia in range(s["natome"]):
        nb, nl = int(s["nbord"][ia]), int(s["nlmsa"][ia])
        pts = s["xyz"][:3, s["ibord"][:nb, ia] - 1].T - s["posi"][:, ia]
        r = np.maximum(np.linalg.norm(pts, axis=1), 0.5 * s["rmt"][ia])
        ls = np.array([L for L in range(int(s["lmax"][ia]) + 1) for m in range(-L, L + 1)])
        out.append(np.stack([spherical_yn(L, k * r) for L in ls], axis=1) * s["ylmato"][:nb, :nl, ia])
    return out


def atom_tau_fixture():
    """The reference mat (mat.f90:5) on the three-atom cluster: Taull after
    its symmetrisation (real energy; real and complex basis) and, with
    FDM_comp at complex energy, the irregular cluster solution and Tau_RI."""
    a = ATOMS
    posi = np.array(a["posi"]) / BOHR
    s = R.atom_system(a["R"] / BOHR, a["h"] / BOHR, a["lmaxso"], posi, np.array(a["rmt"]) / BOHR, a["lmax"])
    V = np.random.default_rng(7).uniform(-0.3, 0.3, s["npoint"]) - 0.8
    out = dict(V=V)
    E = 0.7
    phi = atom_phiato(s, E, 0.0, False)
    for basereel in (True, False):
        T, _ = R.mat_tau(s, a["h"] / BOHR, V, E, 0.0, False, False, basereel, phi)
        out[f"taull_basereel{int(basereel)}"] = T
    out["E"], out["phi_real"] = E, np.concatenate([p.reshape(-1) for p in phi])
    Ec, Ei = 0.7, 0.05
    phi = atom_phiato(s, Ec, Ei, True)
    phi2 = atom_phiato_irregular(s, Ec, Ei)
    _, TR = R.mat_tau(s, a["h"] / BOHR, V, Ec, Ei, True, True, False, phi, phi2)
    lay = R.irregular_layout(s)
    bes, neu = R.phiso_sys(s, True, Ec, Ei, Ec)
    xi = R.mat_solve_sys(s, V, Ec, Ei, bes, neu, True, True, False, phi, phiato2=phi2)
    rows = np.arange(0, lay["nligne"], ROW_STRIDE)
    out.update(Ec=Ec, Eimag=Ei, phi_c=np.concatenate([p.reshape(-1) for p in phi]),
               phi2_c=np.concatenate([p.reshape(-1) for p in phi2]), tau_ri=TR, irr_rows=rows, x_irr=xi[rows],
               irr_nligne=lay["nligne"], irr_ianew=lay["ianew"], irr_lb1=lay["lb1"], irr_lb2=lay["lb2"],
               irr_newinv=lay["newinv"], irr_new=lay["new"])
    np.savez_compressed(os.path.join(HERE, "atoms_tau_R1.9_h0.25.npz"), **out)
    print("atom tau", lay["nligne"], np.abs(TR).max())


def rel_fixtures():
    """mat_solve with Relativiste = T: the P1 cluster of 1.2 A and the
    three-atom cluster, a random potential and gradient."""
    Ra, h, lm = 1.2, 0.25, 2
    cl = O.Cluster(Ra / BOHR, h / BOHR, lm)
    cgrad = R.grid(Ra / BOHR, h / BOHR)["cgrad"]
    rng = np.random.default_rng(21)
    V = rng.uniform(-0.5, 0.5, cl.npoint) - 1.0
    G = rng.normal(size=(cl.npoint, 3)) * 40.0
    out = dict(R=Ra, h=h, lmaxso=lm, V=V, gradvr=G, cgrad=cgrad)
    for b, (E, Ei, E_comp, cal, basereel) in enumerate(ATOM_BRANCHES):
        bes, neu = R.phiso(cl, E_comp, E, Ei, E)
        out[f"x{b}"] = R.mat_solve(cl, V, E, Ei, bes, neu, E_comp, cal, basereel, gradvr=G, cgrad=cgrad)
    np.savez_compressed(os.path.join(HERE, "rel_R1.2_h0.25.npz"), **out)
    a = ATOMS
    posi = np.array(a["posi"]) / BOHR
    s = R.atom_system(a["R"] / BOHR, a["h"] / BOHR, a["lmaxso"], posi, np.array(a["rmt"]) / BOHR, a["lmax"])
    rng = np.random.default_rng(22)
    V = rng.uniform(-0.3, 0.3, s["npoint"]) - 0.8
    G = rng.normal(size=(s["npoint"], 3)) * 40.0
    rows = np.arange(0, s["nligne"], ROW_STRIDE)
    out = dict(V=V, gradvr=G, rows=rows)
    for b, (E, Ei, E_comp, cal, basereel) in enumerate(ATOM_BRANCHES):
        bes, neu = R.phiso_sys(s, E_comp, E, Ei, E)
        x = R.mat_solve_sys(s, V, E, Ei, bes, neu, E_comp, cal, basereel, atom_phiato(s, E, Ei, E_comp), gradvr=G)
        out[f"x{b}"] = x[rows]
    np.savez_compressed(os.path.join(HERE, "atoms_rel_R1.9_h0.25.npz"), **out)
    print("relativistic", cl.nligne, s["nligne"])


def ylm_fixture():
    """mat_solve with Ylm_comp = T (complex harmonics in the outer sphere)
    on the P1 cluster of 1.2 A, lmaxso 3, three branch sets."""
    Ra, h, lm = 1.2, 0.25, 3
    cl = O.Cluster(Ra / BOHR, h / BOHR, lm)
    V = np.random.default_rng(23).uniform(-0.5, 0.5, cl.npoint) - 1.0
    out = dict(R=Ra, h=h, lmaxso=lm, V=V)
    for b, (E, Ei, E_comp, basereel) in enumerate(YLM_BRANCHES):
        bes, neu = R.phiso(cl, E_comp, E, Ei, E)
        out[f"x{b}"] = R.mat_solve(cl, V, E, Ei, bes, neu, E_comp, True, basereel, ylm_comp=True)
    np.savez_compressed(os.path.join(HERE, "ylmc_R1.2_h0.25.npz"), **out)
    print("ylm_comp", cl.nligne)


YLM_BRANCHES = [(0.7, 0.05, True, False), (0.7, 0.0, False, False), (0.7, 0.0, False, True)]


def sphere_inputs(s, ia):
    """A radial grid (geometric from 1e-4 bohr, ratio 1.03, to 1.3 x the
    farthest border point) and a screened Coulomb potential for atom ia."""
    nb = int(s["nbord"][ia])
    bx = s["xyz"][:3, s["ibord"][:nb, ia] - 1].T
    dmax = np.linalg.norm(bx - s["posi"][:, ia], axis=1).max()
    r = [1e-4]
    while r[-1] < 1.3 * dmax:
        r.append(r[-1] * 1.03)
    r = np.array(r)
    Z = (26, 14, 8)[ia]
    return r, -2 * Z / r * np.exp(-r / 0.4) - 0.5, Z


SPHERE_CASES = [(0.7, 0.05, False), (0.9, 0.0, False), (0.7, 0.05, True), (0.9, 0.0, True)]


def sphere_fixture():
    """The reference's Sphere (sphere.f90:8) for each atom of the three-atom
    cluster: phiato at its border points and the atomic Tau, complex and
    real energy, with and without the scalar-relativistic terms; and with
    n_RI = 2 the irregular solution's phiato."""
    a = ATOMS
    posi = np.array(a["posi"]) / BOHR
    s = R.atom_system(a["R"] / BOHR, a["h"] / BOHR, a["lmaxso"], posi, np.array(a["rmt"]) / BOHR, a["lmax"])
    out = {}
    for ia in range(s["natome"]):
        r, V, Z = sphere_inputs(s, ia)
        out[f"r{ia}"], out[f"V{ia}"], out[f"Z{ia}"] = r, V, Z
        for c, (Ec, Ei, rel) in enumerate(SPHERE_CASES):
            phi, tau = R.sphere(s, ia + 1, r, V, Z, a["rmt"][ia] / BOHR, int(s["lmax"][ia]), rel, Ec, Ei, Ec + 0.3)
            out[f"phi{ia}_{c}"], out[f"tau{ia}_{c}"] = phi, tau
            # the irregular solution's phiato (n_RI = 2)
            _, out[f"irr{ia}_{c}"], _ = R.sphere(s, ia + 1, r, V, Z, a["rmt"][ia] / BOHR, int(s["lmax"][ia]), rel,
                                                 Ec, Ei, Ec + 0.3, n_ri=2)
    np.savez_compressed(os.path.join(HERE, "sphere_R1.9_h0.25.npz"), **out)
    print("sphere", [len(out[f"r{ia}"]) for ia in range(s["natome"])])


# (lmax, Jseuil, Lseuil, nbseuil, point-group operations, Core_resolved, Ecinetic, Eimag, relativistic,
#  rotated frames, atomic number)
TENSOR_CASES = [(2, 1, 0, 1, [1], False, 0.7, 0.05, False, False, 26),
                (2, 1, 0, 1, [1], False, 0.9, 0.0, False, True, 26),
                (3, 3, 1, 1, [1], True, 0.7, 0.05, False, True, 26),
                (3, 2, 1, 2, [1, 22, 23, 24, 25, 40, 41, 42], False, 0.9, 0.02, False, False, 26),
                (2, 1, 0, 1, [1], False, 0.6, 0.03, True, True, 26),
                (2, 1, 0, 1, [1], False, 0.7, 0.05, False, False, 1)]


def tensor_fixture():
    """The reference's Tenseur_car (tensor.f90:15, with its Radial_integral)
    for the absorbing atom of the three-atom cluster (atom 0's radial grid
    and potential): K, L3 (core resolved) and L23 edges, E1E1, E1E2 and E2E2,
    no symmetry and mmm, complex and real energy, rotated frames, a random
    Taull and synthetic core functions."""
    a = ATOMS
    posi = np.array(a["posi"]) / BOHR
    s = R.atom_system(a["R"] / BOHR, a["h"] / BOHR, a["lmaxso"], posi, np.array(a["rmt"]) / BOHR, a["lmax"])
    r, V, _ = sphere_inputs(s, 0)
    rmt = a["rmt"][0] / BOHR
    out = dict(r=r, V=V, rmt=rmt, rmtsd=1.1 * rmt, ncase=len(TENSOR_CASES))
    rng = np.random.default_rng(11)
    for k, (lmax, js, ls, nb, ops, cr, Ec, Ei, rel, rot, Z) in enumerate(TENSOR_CASES):
        nlm = (lmax + 1) ** 2
        tl = rng.normal(size=(nlm, nlm)) + 1j * rng.normal(size=(nlm, nlm))
        psii = np.stack([r * np.exp(-26 * r / (1 + j)) * (1 + 0.3 * j * r) for j in range(nb)])
        Es = np.array([520.0, 510.0][:nb])
        ri = np.linalg.qr(rng.normal(size=(3, 3)))[0] if rot else np.eye(3)
        ra = np.linalg.qr(rng.normal(size=(3, 3)))[0] if rot else np.eye(3)
        tabs = R.tensor_tables(ops)
        dd, dq, qq = R.tenseur_car(r, V, Z, rmt, 1.1 * rmt, lmax, Ec, Ei, Ec + 0.3, Ec, Es, psii, tl, js, ls, tabs,
                                   core_resolved=cr, relativiste=rel, rot_int=ri, rot_atom_abs=ra)
        c = f"c{k}_"
        out.update({c + "case": np.array([lmax, js, ls, nb, int(cr), int(rel), Z]), c + "E": np.array([Ec, Ei]),
                    c + "Eseuil": Es, c + "psii": psii, c + "taull": tl, c + "rot_int": ri, c + "rot_atom_abs": ra,
                    c + "dd": dd, c + "dq": dq, c + "qq": qq})
        out.update({c + n: v for n, v in tabs.items()})
    np.savez_compressed(os.path.join(HERE, "tensor_R1.9.npz"), **out)
    print("tensor cases", len(TENSOR_CASES))


def main():
    if not R.available():
        raise SystemExit("oracle/_ref/libfdmnes_ref.so is missing: make -f oracle/Makefile.ref")
    for name, Ra, h, lm, E, Ei, E_comp, cal, basereel, seed in CASES:
        cl = O.Cluster(Ra / BOHR, h / BOHR, lm)
        rng = np.random.default_rng(seed)
        V = rng.uniform(-0.5, 0.5, cl.npoint) - 1.0
        lb1, lb2, newinv, newp = R.newind(cl)
        bes, neu = R.phiso(cl, E_comp, E, Ei, E)
        x = R.mat_solve(cl, V, E, Ei, bes, neu, E_comp, cal, basereel)
        np.savez_compressed(
            os.path.join(HERE, f"{name}.npz"), R=Ra, h=h, lmaxso=lm, E=E, Eimag=Ei, E_comp=E_comp, cal_comp=cal,
            basereel=basereel, V=V, lb1=lb1, lb2=lb2, newinv=newinv, new=newp, bessel=bes, neuman=neu, x=x)
        print(name, cl.nligne, np.abs(x).max())
    # special functions
    zs = [0.3 + 0.05j, 1.7 + 0.2j, 4.2 + 0.01j, 9.5 + 1.0j]
    fn = [np.sqrt((z / 2.0) / np.pi) for z in zs]
    bc = [R.cbessneu(f, z, 6) for f, z in zip(fn, zs)]
    zr = [0.4, 1.9, 5.5, 12.0]
    br = [R.cbessneur(np.sqrt(z / 3.0 / np.pi), z, 6) for z in zr]
    dirs = np.array([[1.0, 2.0, 3.0], [0.0, 0.0, 2.0], [0.0, 0.0, -1.5], [-1.0, 0.5, 0.0], [0.3, -0.7, -0.2]])
    ylm = [R.ylm_real(4, v, float(np.linalg.norm(v))) for v in dirs]
    np.savez_compressed(os.path.join(HERE, "special_functions.npz"), zs=np.array(zs), fn=np.array(fn),
                        bes_c=np.array([b for b, _ in bc]), neu_c=np.array([n for _, n in bc]), zr=np.array(zr),
                        fnr=np.array([np.sqrt(z / 3.0 / np.pi) for z in zr]),
                        bes_r=np.array([b for b, _ in br]), neu_r=np.array([n for _, n in br]), dirs=dirs,
                        ylm=np.array(ylm))
    print("special functions written")
    grid_fixtures()
    atom_fixture()
    atom_tau_fixture()
    rel_fixtures()
    ylm_fixture()
    sphere_fixture()
    tensor_fixture()


if __name__ == "__main__":
    main()
