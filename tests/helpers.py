"""Shared builders for the tests: small synthetic systems assembled by the
product's cluster builder, with the oracle's solution beside them."""
from __future__ import annotations

import numpy as np

BOHR = 0.52917721067
RYDB = 13.605693012905


def small_system(R=1.6, h=0.25, lmaxso=3, E=0.6, eimag=0.0367, seed=0, E_comp=True, basereel=False,
                 source=True):
    """Returns dict(cluster, lb1, lb2, rp, col, vals, rhs, V, pts, E, eimag)."""
    from fdmnes_amd import cluster as CL
    from fdmnes_amd import configs as cfg

    cl = CL.Cluster(R / BOHR, h / BOHR, lmaxso)
    pts = cl.points
    rng = np.random.default_rng(seed)
    V = rng.uniform(-0.3, 0.3, cl.npoint) - 0.8
    Eimag = eimag if E_comp else 0.0
    rp, col, vals, b_inc = cl.assemble(V, E, Eimag, cfg.ecinetic_out(E), E_comp, basereel)
    lb1, lb2 = cl.profile()
    cal = E_comp or not basereel
    if source:
        w = cfg.FE_FP64
        s = cfg.sources(w, pts, cl.nligne)
        rhs = s.astype(np.complex128) if cal else s
    else:
        rhs = b_inc
    return dict(cluster=cl, lb1=lb1, lb2=lb2, rp=rp, col=col, vals=vals, rhs=rhs, V=V, pts=pts, E=E,
                eimag=Eimag, cal=cal)


def oracle_solution(sys):
    from oracle import oracle as O
    st, x, fr = O.mat_solve_csr(sys["lb1"], sys["lb2"], sys["rp"], sys["col"], sys["vals"], sys["rhs"], sys["cal"])
    assert st == 0, fr
    return x


def dense_matrix(sys):
    n = len(sys["lb1"])
    A = np.zeros((n, n), complex)
    rp, col, vals = sys["rp"], sys["col"], sys["vals"]
    for i in range(n):
        A[i, col[rp[i]:rp[i + 1]] - 1] = vals[rp[i]:rp[i + 1]]
    return A
