"""Symmetry-reduced systems of the subgroups of mmm other than mmm itself on
the GPU: the mirror groups m (two representations, wedge z >= 0) and mm2
(four, wedge x, y >= 0), and the groups with rotations or the inversion, -1,
2 about z (two each) and 2/m, 222 (four each), on posequiv's calculation
areas.  Every representation at two energies in one grouped solve, in FP64
and with FP32 factors and FP64 refinement, reproduces the oracle's
whole-cluster solution for a potential with exactly the group's symmetry;
and the representation loop (device assembly of each representation's rows
from its tables, grouped solves) does the same for the groups with rotations
or the inversion.  Bound 1e-10, the solver's parity bar."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.test_cluster_mmm import BOHR, _irreps, _potential
from tests.test_gpu_mmm import _solve_reduced

pytestmark = pytest.mark.gpu


def _whole(R, h, lm, group, ens):
    o = O.Cluster(R, h, lm)
    xyz = o.xyz[:o.npoint, :3]
    row = {tuple(k): i for i, k in enumerate(np.round(xyz / h).astype(int))}
    full = {}
    for e in ens:
        st, xf, *_ = o.solve(_potential(xyz, group), *e, True, True, False)
        assert st == 0
        full[e] = xf
    return row, o.nligne - o.nlmso, full


def _rel(c, x, xf, row, nsof, h):
    hm = c.harmonics
    rows = [row[tuple(k)] for k in np.round(c.points / h).astype(int)]
    return max(np.abs(x[:c.npoint, :c.nlmso] / c.rvol[:, None] - xf[rows][:, hm]).max(),
               np.abs(x[c.npoint:, :c.nlmso] - xf[nsof + hm][:, hm]).max()) / np.abs(xf).max()


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("group", [3, 6, 2, 4, 5, 7])
def test_grouped_group_representations_match_oracle(group, precision):
    from fdmnes_amd import cluster as CL
    R, h, lm = 2.0 / BOHR, 0.25 / BOHR, 3
    reps = [CL.Cluster(R, h, lm, irrep=g, group=group) for g in _irreps(group)]
    ens = [(0.35, 0.0367, 0.55), (0.8, 0.0367, 1.0)]
    print(f"group {group} precision {precision}: {len(reps)} representations x {len(ens)} energies, rows",
          [c.nligne for c in reps], flush=True)
    out = _solve_reduced(reps, lambda p: _potential(p, group), ens, True, False, precision)
    row, nsof, full = _whole(R, h, lm, group, ens)
    for c, e, x in out:
        d = _rel(c, x, full[e], row, nsof, h)
        print("irrep", c.irrep, "energy", e[0], "rel", d, flush=True)
        assert d <= 1e-10


@pytest.mark.parametrize("group", [2, 4, 5, 7])
def test_representation_loop_of_rotation_groups_matches_oracle(group):
    from fdmnes_amd import cluster as CL
    from fdmnes_amd.energy_loop import RepresentationLoop
    R, h, lm = 2.0 / BOHR, 0.25 / BOHR, 3
    reps = [CL.Cluster(R, h, lm, irrep=g, group=group) for g in _irreps(group)]
    ens = [(0.15 + 0.25 * k, 0.0367, 0.45 + 0.25 * k) for k in range(3)]
    print(f"group {group}: representation loop over {len(reps)} representations x {len(ens)} energies", flush=True)
    out = RepresentationLoop(reps, precision=0).run([_potential(c.points, group) for c in reps], ens,
                                                    keep_solutions=True)
    row, nsof, full = _whole(R, h, lm, group, ens)
    for k, e in enumerate(ens):
        for r, c in enumerate(reps):
            d = _rel(c, out["x"][k][r], full[e], row, nsof, h)
            print("energy", k, "irrep", c.irrep, "rel", d, flush=True)
            assert d <= 1e-10
