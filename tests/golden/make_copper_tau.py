"""Writes tests/golden/cu13_real_tau.npz: for the real-arithmetic copper
cluster (fcc Cu, R = 3.0 A, h = 0.24 A, l <= 2; Basereel, no complex energy)
at four of its 91 energies (-5, 10, 25 and 40 eV), the oracle's solution
with calcMatRow's incoming-wave right-hand sides reduced to tau in the real
basis over the block of the nlmso innermost grid rows (mat.f90:297-405:
normalisation I + i X of the outer rows, invcomp, amplitudes, -i a^T conj(a)),
and sigma_tau = -Im sum_L tau_LL.  Beside them the same reduction of a
partial-pivoting dense solve (numpy) of the oracle's dense matrix, which
shows how far the reference algorithm itself is from a backward-stable
solve on this undamped system.  The oracle reproduces the flang-built
reference bit for bit on this branch.  About a minute per energy.

    python tests/golden/make_copper_tau.py
"""
from __future__ import annotations

import multiprocessing as mp
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
ENERGY_INDEX = [0, 30, 60, 90]


def one(k):
    sys.path.insert(0, ROOT)
    from fdmnes_amd import configs as cfg
    from oracle import oracle as O
    w = cfg.CU_REAL
    cl = O.Cluster(w.R / cfg.BOHR, w.h / cfg.BOHR, w.lmaxso)
    V = cfg.potential(w, cl.xyz[: cl.npoint, :3])
    E = float(w.energies_Ry[k])
    nso = cl.nligne - cl.nlmso
    st, x, fr, _, _, _ = cl.solve(V, E, 0.0, cfg.ecinetic_out(E), False, False, True)
    assert st == 0, fr
    A, b = cl.assemble_dense(V, E, 0.0, cfg.ecinetic_out(E), False, False, True)
    xd = np.linalg.solve(A, b)
    tau = O.tau(x, nso, True, False, [(0, cl.nlmso)])[0][0]
    taud = O.tau(xd, nso, True, False, [(0, cl.nlmso)])[0][0]
    res = np.linalg.norm(b - A @ x, axis=0) / np.linalg.norm(b, axis=0)
    return E, tau, taud, x[nso:], res.max()


def main():
    with mp.get_context("spawn").Pool(len(ENERGY_INDEX)) as pool:
        out = pool.map(one, ENERGY_INDEX)
    E = np.array([o[0] for o in out])
    tau = np.stack([o[1] for o in out])
    taud = np.stack([o[2] for o in out])
    np.savez_compressed(os.path.join(HERE, "cu13_real_tau.npz"), energy_index=np.array(ENERGY_INDEX), E=E,
                        tau=tau, sigma_tau=-np.imag(np.trace(tau, axis1=1, axis2=2)), tau_dense=taud,
                        sigma_tau_dense=-np.imag(np.trace(taud, axis1=1, axis2=2)),
                        x_outer=np.stack([o[3] for o in out]), residual=np.array([o[4] for o in out]))


if __name__ == "__main__":
    main()
