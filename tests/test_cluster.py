"""The product's cluster builder (host C++) against the oracle: same radial
order, neighbour structure, profile and compressed rows; and the workload
definitions."""
import numpy as np
import pytest

from oracle import oracle as O

BOHR = 0.52917721067


@pytest.mark.parametrize("R,h,lm", [(1.2, 0.25, 2), (2.0, 0.2, 3), (3.0, 0.24, 2)])
def test_builder_matches_oracle(R, h, lm):
    from fdmnes_amd import cluster as CL
    c = CL.Cluster(R / BOHR, h / BOHR, lm)
    o = O.Cluster(R / BOHR, h / BOHR, lm)
    assert (c.npoint, c.npso, c.nligne, c.nsortf, c.nsort) == (o.npoint, o.npso, o.nligne, o.nsortf, o.nsort)
    lb1, lb2 = c.profile()
    ol1, ol2, _, _ = o.profile()
    assert np.array_equal(lb1, ol1) and np.array_equal(lb2, ol2)
    assert np.array_equal(c.points, o.xyz[: c.npoint, :3])
    V = np.random.default_rng(1).uniform(-1, 0, c.npoint)
    for E_comp, basereel in [(True, False), (False, True), (False, False)]:
        E, Ei = 0.5, (0.03 if E_comp else 0.0)
        cal = E_comp or not basereel
        rp, col, vals, rhs = c.assemble(V, E, Ei, E, E_comp, basereel)
        orp, ocol, ovals, orhs = o.assemble_csr(V, E, Ei, E, E_comp, cal, basereel)
        assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
        assert np.abs(vals - ovals).max() <= 1e-14 * np.abs(ovals).max()
        assert np.abs(rhs - orhs).max() <= 1e-13 * np.abs(orhs).max()


def test_workload_sizes():
    from fdmnes_amd import cluster as CL
    from fdmnes_amd import configs as cfg
    assert len(cfg.lattice_sites("fcc", 3.61, cfg.CU_REAL.atoms_radius)) == 13
    assert len(cfg.lattice_sites("bcc", 2.866, 5.0)) == 59
    c = CL.Cluster(cfg.CU_REAL.R / BOHR, cfg.CU_REAL.h / BOHR, cfg.CU_REAL.lmaxso)
    assert c.npoint == 8217 and c.nligne == 8226


def test_sources_and_potential():
    from fdmnes_amd import configs as cfg
    pts = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.5, -0.5]])
    V = cfg.potential(cfg.FE_FP64, pts)
    assert np.all(np.isfinite(V)) and V[0] < -100  # absorber well, regularised at h/2
    s = cfg.sources(cfg.FE_FP64, pts, 5)
    assert s.shape == (5, 16) and np.all(s[3:] == 0)
    assert abs(s[0, 0] - 1 / np.sqrt(4 * np.pi)) < 1e-15
    # the real harmonics agree with the oracle's (reference convention)
    y = cfg.real_ylm(3, pts[1:])
    for p, yy in zip(pts[1:], y):
        assert np.abs(yy - O.ylm_real(3, p, float(np.linalg.norm(p)))).max() < 1e-14
