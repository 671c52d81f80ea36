"""The atoms' radial solutions and phiato (Sphere, sphere.f90:8-210): the
oracle's restatement (orc_sphere) against the reference's own Sphere
(tests/golden/sphere_R1.9_h0.25.npz, written by make_golden.py from
oracle/_ref: the three atoms of the atom fixture, complex and real energy,
with and without the scalar-relativistic terms) within 1e-12 (the complex
square roots, exponentials and divisions differ in their last bits), and
against what the mathematics fixes for a constant potential (the regular
solution is j_L(q r) with q^2 = Enervide - V0, so phiato / Ylm follows
j_L(q r) from point to point and Tau = -W(j_L(q r), j_L(k r)) / W(j_L(q r),
-i h_L(k r)) at the matching radius), to the radial grid's discretisation."""
import numpy as np
import pytest

from oracle import oracle as O

from .test_oracle_atoms import BOHR, G

SPHERE = G.replace("atoms_R1.9_h0.25.npz", "sphere_R1.9_h0.25.npz")
CASES = [(0.7, 0.05, False), (0.9, 0.0, False), (0.7, 0.05, True), (0.9, 0.0, True)]


def _cluster():
    g = dict(np.load(G))
    o = O.Cluster(float(g["R"]) / BOHR, float(g["h"]) / BOHR, int(g["lmaxso"]),
                  atoms=(g["posi"] / BOHR, g["rmt"] / BOHR, g["lmax"]))
    return g, o


def border(o, g, ia):
    """The atom's border points (bohr), its position, its harmonics there."""
    d = o.atom(ia + 1)
    return o.xyz[d["ibord"] - 1, :3], g["posi"][ia] / BOHR, d["ylmato"]


@pytest.mark.parametrize("case", range(4))
@pytest.mark.parametrize("ia", range(3))
def test_oracle_sphere_matches_reference(ia, case):
    g, o = _cluster()
    f = np.load(SPHERE)
    Ec, Ei, rel = CASES[case]
    bx, pos, ylm = border(o, g, ia)
    lmax = int(g["lmax"][ia])
    phi, tau, cx = O.sphere(f[f"r{ia}"], f[f"V{ia}"], int(f[f"Z{ia}"]), g["rmt"][ia] / BOHR, lmax, rel, Ec, Ei,
                            Ec + 0.3, bx, pos, ylm)
    assert cx == (Ei != 0)
    ref, tref = f[f"phi{ia}_{case}"], f[f"tau{ia}_{case}"]
    assert np.abs(phi - ref).max() <= 1e-12 * np.abs(ref).max()
    t = np.array([tref[L * L] for L in range(lmax + 1)])  # Tau_ato's diagonal, one value per L
    assert np.abs(tau - t).max() <= 1e-12 * np.abs(t).max()
    for L in range(lmax + 1):  # the same Tau for every m of L
        assert np.all(tref[L * L:(L + 1) ** 2] == tref[L * L])


def test_constant_potential_closed_forms():
    from scipy.special import spherical_jn, spherical_yn
    g, o = _cluster()
    ia = 0
    bx, pos, ylm = border(o, g, ia)
    rel = np.linalg.norm(bx - pos, axis=1)
    r = [1e-4]
    while r[-1] < 1.2 * rel.max():
        r.append(r[-1] * 1.002)
    r = np.array(r)
    V0, Ev, Ec, rmt, lmax = -1.1, 0.4, 0.7, g["rmt"][ia] / BOHR, 2
    phi, tau, cx = O.sphere(r, np.full(len(r), V0), 0, rmt, lmax, False, Ec, 0.0, Ev, bx, pos, ylm)
    assert not cx
    q, k = np.sqrt(Ev - V0), np.sqrt(Ec)
    # the point of the matching (the first radial point above rmtg)
    rm = rmt
    for L in range(lmax + 1):
        sel = (np.abs(ylm[:, L * L + L]) > 0.2) & (rel > 0.5 * rmt)
        ratio = phi.real[sel, L * L + L] / ylm[sel, L * L + L] / spherical_jn(L, q * rel[sel])
        assert np.ptp(ratio) <= 2e-4 * np.abs(ratio).max(), L
        j = spherical_jn(L, q * rm)
        dj = q * spherical_jn(L, q * rm, derivative=True)
        jk, djk = spherical_jn(L, k * rm), k * spherical_jn(L, k * rm, derivative=True)
        hk = jk + 1j * spherical_yn(L, k * rm)
        dhk = djk + 1j * k * spherical_yn(L, k * rm, derivative=True)
        we = j * djk - dj * jk
        ws = j * (-1j * dhk) - dj * (-1j * hk)
        t = -we / ws
        assert abs(tau[L] - t) <= 2e-3 * abs(t), (L, tau[L], t)


@pytest.mark.parametrize("case", range(4))
@pytest.mark.parametrize("ia", range(3))
def test_oracle_irregular_solution_matches_reference(ia, case):
    """The irregular solution's phiato (Sch_radial_solsing, sphere.f90:2591,
    Sphere with n_RI = 2) against the reference's own Sphere within 1e-12;
    the regular phiato of the same call unchanged."""
    g, o = _cluster()
    f = np.load(SPHERE)
    Ec, Ei, rel = CASES[case]
    bx, pos, ylm = border(o, g, ia)
    lmax = int(g["lmax"][ia])
    phi, phi2, tau, cx = O.sphere_ri(f[f"r{ia}"], f[f"V{ia}"], int(f[f"Z{ia}"]), g["rmt"][ia] / BOHR, lmax, rel, Ec,
                                     Ei, Ec + 0.3, bx, pos, ylm)
    ref = f[f"irr{ia}_{case}"]
    assert np.abs(phi2 - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.abs(phi - f[f"phi{ia}_{case}"]).max() <= 1e-12 * np.abs(f[f"phi{ia}_{case}"]).max()
    if Ei == 0:  # real radial solutions: the irregular phiato keeps only its real part
        assert np.all(phi2.imag == 0)


def test_irregular_solution_in_a_constant_potential_is_the_outgoing_wave():
    """With the same kinetic energy inside the sphere as outside (V0 =
    Enervide - Ecinetic) the inward integration continues -i r h_L(k r):
    phiato2 / Ylm follows -i fnorm h_L(k r) with fnorm = sqrt(k / pi), to the
    radial grid's discretisation."""
    from scipy.special import spherical_jn, spherical_yn
    g, o = _cluster()
    ia = 0
    bx, pos, ylm = border(o, g, ia)
    rel = np.linalg.norm(bx - pos, axis=1)
    r = [1e-4]
    while r[-1] < 1.2 * rel.max():
        r.append(r[-1] * 1.002)
    r = np.array(r)
    Ec, Ei, Ev = 0.7, 0.05, 0.4
    V0 = Ev - Ec
    rmt, lmax = g["rmt"][ia] / BOHR, 2
    _, phi2, _, cx = O.sphere_ri(r, np.full(len(r), V0), 0, rmt, lmax, False, Ec, Ei, Ev, bx, pos, ylm)
    assert cx
    k = np.sqrt(Ec + 1j * Ei)
    fn = np.sqrt(k / np.pi)
    for L in range(lmax + 1):
        sel = (np.abs(ylm[:, L * L + L]) > 0.2) & (rel > 0.5 * rmt)
        want = -1j * fn * (spherical_jn(L, k * rel[sel]) + 1j * spherical_yn(L, k * rel[sel]))
        got = phi2[sel, L * L + L] / ylm[sel, L * L + L]
        assert np.abs(got - want).max() <= 2e-4 * np.abs(want).max(), L
