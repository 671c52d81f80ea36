"""Speed and results of the oracle's mat_solve restatement against the
flang-built reference (oracle/_ref) on the real-arithmetic copper cluster
(fcc Cu, R = 3.0 A, h = 0.24 A, l <= 2: 8,217 unknowns, 9 right-hand sides,
Basereel without complex energy, the reference's default branch).  Both run
single-threaded in this process, each assembling its rows with its own
calcMatRow.  Needs the reference sources (to build oracle/_ref); run from the
repository root:

    python -m oracle.time_vs_ref
"""
from __future__ import annotations

import time

import numpy as np

from fdmnes_amd import cluster as CL
from fdmnes_amd import configs as cfg
from oracle import oracle as O
from oracle import ref as R


def main():
    if not R.available():
        raise SystemExit("oracle/_ref is not built (make -f oracle/Makefile.ref)")
    w = cfg.CU_REAL
    pcl = CL.Cluster(w.R / cfg.BOHR, w.h / cfg.BOHR, w.lmaxso)
    cl = O.Cluster(w.R / cfg.BOHR, w.h / cfg.BOHR, w.lmaxso)
    V = cfg.potential(w, pcl.points)
    E = float(w.energies_Ry[len(w.energies_Ry) // 2])
    print(f"cluster {w.name}: nligne={cl.nligne}, nlmso={cl.nlmso}, E = {E * cfg.RYDB:.2f} eV")
    t0 = time.time()
    br, nr = R.phiso(cl, False, E, 0.0, E)
    xr = R.mat_solve(cl, V, E, 0.0, br, nr, False, False, True)
    t_ref = time.time() - t0
    t0 = time.time()
    st, xo, fr, tf, tt, nnz = cl.solve(V, E, 0.0, E, False, False, True)
    t_port = time.time() - t0
    assert st == 0
    rel = np.abs(xo - xr).max() / np.abs(xr).max()
    print(f"reference mat_solve {t_ref:.2f} s, oracle {t_port:.2f} s, oracle/reference time {t_port / t_ref:.3f}, "
          f"max |x_oracle - x_ref| / max |x_ref| = {rel:.3e}, bit-identical: {np.array_equal(xo, xr)}")


if __name__ == "__main__":
    main()
