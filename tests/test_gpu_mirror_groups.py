"""Symmetry-reduced systems of the mirror groups m (two representations,
wedge z >= 0) and mm2 (four, wedge x, y >= 0) on the GPU: every
representation at two energies in one grouped solve (FP32 factors with FP64
refinement) reproduces the oracle's whole-cluster solution for a potential
with exactly the group's mirrors.  Bound 1e-10, the solver's parity bar."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.test_cluster_mmm import BOHR, _irreps, _potential
from tests.test_gpu_mmm import _solve_reduced

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("group", [3, 6])
def test_grouped_mirror_group_representations_match_oracle(group):
    from fdmnes_amd import cluster as CL
    R, h, lm = 2.0 / BOHR, 0.25 / BOHR, 3
    reps = [CL.Cluster(R, h, lm, irrep=g, group=group) for g in _irreps(group)]
    ens = [(0.35, 0.0367, 0.55), (0.8, 0.0367, 1.0)]
    print(f"group {group}: {len(reps)} representations x {len(ens)} energies, rows",
          [c.nligne for c in reps], flush=True)
    out = _solve_reduced(reps, lambda p: _potential(p, group), ens, True, False, 1)
    o = O.Cluster(R, h, lm)
    xyz = o.xyz[:o.npoint, :3]
    row = {tuple(k): i for i, k in enumerate(np.round(xyz / h).astype(int))}
    nsof = o.nligne - o.nlmso
    full = {}
    for e in ens:
        st, xf, *_ = o.solve(_potential(xyz, group), *e, True, True, False)
        assert st == 0
        full[e] = xf
    for c, e, x in out:
        xf = full[e]
        hm = c.harmonics
        rows = [row[tuple(k)] for k in np.round(c.points / h).astype(int)]
        d = max(np.abs(x[:c.npoint, :c.nlmso] / c.rvol[:, None] - xf[rows][:, hm]).max(),
                np.abs(x[c.npoint:, :c.nlmso] - xf[nsof + hm][:, hm]).max()) / np.abs(xf).max()
        print("irrep", c.irrep, "energy", e[0], "rel", d, flush=True)
        assert d <= 1e-10
