"""The oracle against the reference's own outputs (golden vectors written by
tests/golden/make_golden.py from the flang-built reference) and, where
oracle/_ref is present, against the reference live."""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
# the mat_solve cases (the grid builders' fixtures, grid_*, are tests/test_oracle_grid.py's,
# the atom cluster's, atoms_*, tests/test_oracle_atoms.py's)
GOLD = sorted(p for p in glob.glob(os.path.join(HERE, "golden", "*_R*.npz"))
              if not os.path.basename(p).startswith(("grid_", "atoms_", "rel_", "ylmc_", "sphere_", "tensor_")))
BOHR = 0.52917721067


@pytest.mark.parametrize("path", GOLD, ids=[os.path.basename(p)[:-4] for p in GOLD])
def test_oracle_matches_reference_golden(path):
    g = np.load(path)
    cl = O.Cluster(float(g["R"]) / BOHR, float(g["h"]) / BOHR, int(g["lmaxso"]))
    lb1, lb2, newinv, newp = cl.profile()
    assert np.array_equal(lb1, g["lb1"]) and np.array_equal(lb2, g["lb2"])
    assert np.array_equal(newinv, g["newinv"]) and np.array_equal(newp, g["new"])
    E_comp, cal, basereel = bool(g["E_comp"]), bool(g["cal_comp"]), bool(g["basereel"])
    E, Ei = float(g["E"]), float(g["Eimag"])
    bes, neu = cl.phiso(E_comp, E, Ei)
    tol_f = 1e-14 if E_comp else 0.0  # complex libm functions may differ by an ulp
    assert np.abs(bes - g["bessel"]).max() <= tol_f * np.abs(g["bessel"]).max()
    assert np.abs(neu - g["neuman"]).max() <= tol_f * np.abs(g["neuman"]).max()
    st, x, fr, *_ = cl.solve(g["V"], E, Ei, E, E_comp, cal, basereel)
    assert st == 0
    rel = np.abs(x - g["x"]).max() / np.abs(g["x"]).max()
    # the real branches reproduce the reference bit for bit; the complex-energy
    # branch inherits the ulp differences of csin/ccos
    assert rel <= (1e-13 if E_comp else 0.0), rel


def test_special_functions_golden():
    g = np.load(os.path.join(HERE, "golden", "special_functions.npz"))
    for z, f, b, n in zip(g["zs"], g["fn"], g["bes_c"], g["neu_c"]):
        bo, no = O.cbessneu(complex(f), complex(z), 6)
        # the upward recurrence amplifies the ulp differences of csin/ccos for
        # L above |z|; compare where it is stable
        amp = 1.0
        for L in range(7):
            if L:
                amp *= max(1.0, (2 * L - 1) / abs(z)) ** 2
            tol = 1e-13 * amp
            assert abs(bo[L] - b[L]) <= tol * abs(b[L]) + 1e-300, (z, L)
            assert abs(no[L] - n[L]) <= 1e-13 * abs(n[L]), (z, L)
    for z, f, b, n in zip(g["zr"], g["fnr"], g["bes_r"], g["neu_r"]):
        bo, no = O.cbessneur(float(f), float(z), 6)
        assert np.abs(bo - b).max() <= 1e-15 * np.abs(b).max()
        assert np.abs(no - n).max() <= 1e-15 * np.abs(n).max()
    for v, y in zip(g["dirs"], g["ylm"]):
        yo = O.ylm_real(4, v, float(np.linalg.norm(v)))
        assert np.array_equal(yo, y)  # the oracle divides as the reference build does (cylm)


@pytest.mark.ref
def test_oracle_matches_reference_live():
    from oracle import ref as R
    if not R.available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    cl = O.Cluster(1.4 / BOHR, 0.27 / BOHR, 3)
    rng = np.random.default_rng(11)
    V = rng.uniform(-0.5, 0.5, cl.npoint) - 0.9
    a = cl.profile()
    b = R.newind(cl)
    for u, w in zip(a, b):
        assert np.array_equal(u, w)
    for E_comp, cal, basereel in [(True, True, False), (False, False, True), (False, True, False), (False, True, True)]:
        E, Ei = 1.1, (0.04 if E_comp else 0.0)
        br, nr = R.phiso(cl, E_comp, E, Ei, E)
        xr = R.mat_solve(cl, V, E, Ei, br, nr, E_comp, cal, basereel)
        st, xo, *_ = cl.solve(V, E, Ei, E, E_comp, cal, basereel)
        assert st == 0
        rel = np.abs(xo - xr).max() / np.abs(xr).max()
        assert rel <= (1e-13 if E_comp else 0.0), (E_comp, cal, basereel, rel)
