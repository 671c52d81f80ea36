"""The allocation-free sigma (fdmgpu_solver_sigma_noalloc) against
fdmgpu_solver_sigma: the same value bit for bit, on two calls in a row, the
second with fewer and smaller systems (the shared scratch keeps the first
call's partial sums, which the second must not pick up)."""
import numpy as np
import pytest

from tests.helpers import small_system

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def test_noalloc_sigma_equals_sigma_on_two_calls():
    import torch
    from fdmnes_amd.batch import DeviceSolver
    systems = [small_system(R=2.0, seed=31, E=0.6), small_system(R=1.6, seed=32, E=1.1),
               small_system(R=1.2, seed=33, E=1.9)]
    print("no-alloc sigma: sizes", [len(s["lb1"]) for s in systems], flush=True)
    sv = DeviceSolver(0, True, systems[0]["rhs"].shape[1])
    slots = [sv.add_system(s["lb1"], s["lb2"], s["rp"], s["col"]) for s in systems]
    vals = [_dev(s["vals"].astype(np.complex128)) for s in systems]
    b = [_dev(s["rhs"].astype(np.complex128)) for s in systems]
    x = [torch.zeros_like(t) for t in b]
    sv.solve(slots, vals, b, x)
    out = torch.full((4,), np.nan, dtype=torch.float64, device="cuda")
    for call, pick in enumerate([[0, 1, 2], [2]]):  # the second call: one point, the smallest system
        ss = [slots[k] for k in pick]
        ref = sv.sigma(ss, [b[k] for k in pick], [x[k] for k in pick])
        sv.sigma_noalloc(ss, [b[k] for k in pick], [x[k] for k in pick], out)
        got = out.cpu().numpy()[:len(pick)]
        print("call", call, "sigma", ref, "no-alloc", got, flush=True)
        assert np.array_equal(got, ref), (call, got, ref)
        # and the value is the numpy one for the host copies
        for k, v in zip(pick, got):
            s_h = systems[k]["rhs"]
            x_h = x[k].cpu().numpy()
            assert abs(v + np.imag(np.sum(s_h * x_h))) <= 1e-12 * abs(v)


def test_noalloc_sigma_rejects_more_points_than_its_scratch():
    from fdmnes_amd import _lib
    from fdmnes_amd.batch import DeviceSolver
    import torch
    s = small_system(R=1.0, seed=34, E=0.5)
    sv = DeviceSolver(0, True, s["rhs"].shape[1])
    slot = sv.add_system(s["lb1"], s["lb2"], s["rp"], s["col"])
    b = _dev(s["rhs"].astype(np.complex128))
    out = torch.zeros(33, dtype=torch.float64, device="cuda")
    with pytest.raises(Exception) as e:
        sv.sigma_noalloc([slot] * 33, [b] * 33, [b] * 33, out)
    assert "at most 32" in str(e.value)
