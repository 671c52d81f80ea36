"""Each benchmark configuration's own workload on the GPU, where the oracle is
too slow to run at full size (hours per iron energy point): residuals
||b - A x|| / ||b|| computed on the host in FP64 from the same compressed
rows, agreement between the FP64 and mixed-precision solvers, and the
adjoint against central finite differences.

- config 1: one energy of the iron cluster (bcc Fe, 59 atoms, R = 5.0 A,
  h = 0.2 A, 65,283 unknowns, 16 right-hand sides) through the FP64 drop-in;
- config 2: the three iron clusters (R = 3, 4, 5 A) at four energies each,
  twelve systems of three sizes in one grouped FP64 solve;
- config 3: sigma(E) of the mixed-precision solver against the FP64 solver at
  24 of the 200 energies (every 8th and the last), each within 1e-6;
- config 4: d sigma / d V at full size against central differences
  (dV = 1e-6 Ry, 128 FP64 solves) on 48 random grid points and 16 random
  per-atom shifts at one energy, and the checkpoint/recompute path (three
  energies, two factor slots) against the gradient with every point's
  factors resident."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _csr_residual(rp, col, vals, x, b):
    import scipy.sparse as sp
    n = len(rp) - 1
    A = sp.csr_matrix((vals, col - 1, rp), shape=(n, n))
    return np.linalg.norm(b - A @ x, axis=0) / np.linalg.norm(b, axis=0)


def _iron():
    from fdmnes_amd import cluster as CL
    from fdmnes_amd import configs as cfg
    w = cfg.FE_FP64
    cl = CL.Cluster(w.R / cfg.BOHR, w.h / cfg.BOHR, w.lmaxso)
    return w, cl, cl.profile(), cfg.potential(w, cl.points), cfg.sources(w, cl.points, cl.nligne)


def _free():
    import gc
    import torch
    gc.collect()
    torch.cuda.empty_cache()


def test_config1_fp64_dropin_residual():
    from fdmnes_amd import configs as cfg
    from fdmnes_amd import solver as SV
    w, cl, (lb1, lb2), V, s = _iron()
    E = float(w.energies_Ry[137])
    rp, col, vals, _ = cl.assemble(V, E, w.eimag_Ry, cfg.ecinetic_out(E), True, False)
    b = s.astype(np.complex128)
    print("config 1: nligne", cl.nligne, "E", E * cfg.RYDB, "eV", flush=True)
    x, tf, tt = SV.mat_solve(lb1, lb2, rp, col, vals, b, True)
    res = _csr_residual(rp, col, vals, x, b)
    print("fp64 residual max", res.max(), "seconds", tt, flush=True)
    assert res.max() <= 1e-10
    sig64 = cfg.sigma(b, x)
    x32, sweeps, rel = SV.mat_solve_mixed(lb1, lb2, rp, col, vals, b, True)
    sig32 = cfg.sigma(b, x32)
    print("mixed sweeps", sweeps, "sigma rel diff", abs(sig32 - sig64) / abs(sig64), flush=True)
    assert sweeps >= 0 and _csr_residual(rp, col, vals, x32, b).max() <= 1e-10
    assert abs(sig32 - sig64) <= 1e-10 * abs(sig64)
    SV.finalize()
    _free()


def test_config2_three_clusters_four_energies_grouped():
    import torch
    from fdmnes_amd import batch as B
    from fdmnes_amd import cluster as CL
    from fdmnes_amd import configs as cfg
    systems = []
    for k, w in enumerate(cfg.BATCH):
        cl = CL.Cluster(w.R / cfg.BOHR, w.h / cfg.BOHR, w.lmaxso)
        lb1, lb2 = cl.profile()
        V = cfg.potential(w, cl.points)
        s = cfg.sources(w, cl.points, cl.nligne).astype(np.complex128)
        for j in range(4):
            E = float(w.energies_Ry[(23 * j + 7 * k + 3) % len(w.energies_Ry)])
            rp, col, vals, _ = cl.assemble(V, E, w.eimag_Ry, cfg.ecinetic_out(E), True, False)
            systems.append((lb1, lb2, rp, col, vals.astype(np.complex128), s))
        print(w.name, "nligne", cl.nligne, flush=True)
    sv = B.DeviceSolver(B.FP64, True, 16)
    slots = [sv.add_system(a, b_, r, c) for a, b_, r, c, _, _ in systems]
    vals = [torch.from_numpy(v).cuda() for *_, v, _ in systems]
    src = [torch.from_numpy(s).cuda() for *_, s in systems]
    xs = [torch.zeros_like(t) for t in src]
    sv.solve(slots, vals, src, xs)
    for (lb1, lb2, rp, col, v, s), x in zip(systems, xs):
        r = _csr_residual(rp, col, v, x.cpu().numpy(), s).max()
        print("n", len(lb1), "residual", r, flush=True)
        assert r <= 1e-10
    del sv, vals, src, xs
    _free()


def test_config3_sigma_mixed_vs_fp64_per_energy():
    import torch
    from fdmnes_amd import batch as B
    from fdmnes_amd import configs as cfg
    w, cl, (lb1, lb2), V, s = _iron()
    ks = list(range(0, 200, 8)) + [199]  # 26 energies, two groups of 12 and one of 2
    ks = ks[:24]
    src = torch.from_numpy(s.astype(np.complex128)).cuda()
    vals, rp, col = [], None, None
    for k in ks:
        E = float(w.energies_Ry[k])
        rp, col, v, _ = cl.assemble(V, E, w.eimag_Ry, cfg.ecinetic_out(E), True, False)
        vals.append(torch.from_numpy(v.astype(np.complex128)).cuda())
    sig = {}
    for prec in (B.FP64, B.FP32):
        sv = B.DeviceSolver(prec, True, 16)
        slots = [sv.add_system(lb1, lb2, rp, col) for _ in range(12)]
        out = []
        for g0 in range(0, len(ks), 12):
            xs = [torch.zeros_like(src) for _ in range(12)]
            _, sw, _ = sv.solve(slots, vals[g0:g0 + 12], [src] * 12, xs)
            if prec == B.FP32:
                assert all(k >= 0 for k in sw), sw
            out += list(sv.sigma(slots, [src] * 12, xs))
        sig[prec] = np.array(out)
        del sv, xs
        _free()
    rel = np.abs(sig[B.FP32] - sig[B.FP64]) / np.abs(sig[B.FP64])
    for k, r in zip(ks, rel):
        print(f"E = {w.energies_Ry[k] * cfg.RYDB:7.2f} eV  sigma {sig[B.FP64][ks.index(k)]: .12e}  rel err {r:.2e}",
              flush=True)
    assert rel.max() <= 1e-6


def test_config4_adjoint_fd_and_recompute():
    import torch
    from fdmnes_amd import adjoint as AD
    from fdmnes_amd import batch as B
    from fdmnes_amd import configs as cfg
    w, cl, (lb1, lb2), V0, s = _iron()
    npt = cl.npoint
    E = [float(w.energies_Ry[k]) for k in (40, 95, 160)]
    wts = np.array([1.0, -0.7, 0.4])
    # recompute path: three energies, two slots (every backward group refactorises)
    eng = AD.SigmaEngine(cl, s, w.eimag_Ry, slots=2)
    V = torch.tensor(V0, dtype=torch.float64, requires_grad=True)
    (AD.SigmaFunction.apply(V, eng, np.array(E)) * torch.tensor(wts)).sum().backward()
    g_rec = V.grad.numpy().copy()
    print("recomputed", eng.recomputed, flush=True)
    assert eng.recomputed > 0
    del eng
    _free()
    # all resident: three slots
    eng = AD.SigmaEngine(cl, s, w.eimag_Ry, slots=3)
    V = torch.tensor(V0, dtype=torch.float64, requires_grad=True)
    (AD.SigmaFunction.apply(V, eng, np.array(E)) * torch.tensor(wts)).sum().backward()
    g = V.grad.numpy().copy()
    assert eng.recomputed == 0
    del eng
    _free()
    print("recompute vs resident: max diff", np.abs(g_rec - g).max(), "max |g|", np.abs(g).max(), flush=True)
    assert np.abs(g_rec - g).max() <= 1e-12 * np.abs(g).max()
    # central differences of sigma at the middle energy on 48 grid points and
    # 16 atom shifts, against that energy's adjoint gradient
    eng = AD.SigmaEngine(cl, s, w.eimag_Ry, slots=1)
    V = torch.tensor(V0, dtype=torch.float64, requires_grad=True)
    AD.SigmaFunction.apply(V, eng, np.array(E[1:2])).sum().backward()
    g = V.grad.numpy().copy()
    del eng
    _free()
    rng = np.random.default_rng(4)
    pts = rng.choice(npt, 48, replace=False)
    W = cfg.atom_weights(w, cl.points)
    atoms = rng.choice(W.shape[0], 16, replace=False)
    dV = 1e-6
    perturb = [(("V", int(i)), np.eye(1, npt, int(i))[0]) for i in pts] + [(("shift", int(a)), W[a]) for a in atoms]
    src = torch.from_numpy(s.astype(np.complex128)).cuda()
    sv = B.DeviceSolver(B.FP64, True, 16)
    slots = None
    jobs = [(p, sg) for p in range(len(perturb)) for sg in (1, -1)]
    acc = np.zeros((len(perturb), 2))
    for j0 in range(0, len(jobs), 12):
        chunk = jobs[j0:j0 + 12]
        vals = []
        for p, sg in chunk:
            rp, col, v, _ = cl.assemble(V0 + sg * dV * perturb[p][1], E[1], w.eimag_Ry, cfg.ecinetic_out(E[1]), True,
                                        False)
            vals.append(torch.from_numpy(v.astype(np.complex128)).cuda())
        if slots is None:
            slots = [sv.add_system(lb1, lb2, rp, col) for _ in range(12)]
        xs = [torch.zeros_like(src) for _ in chunk]
        sv.solve(slots[:len(chunk)], vals, [src] * len(chunk), xs)
        for (p, sg), val in zip(chunk, sv.sigma(slots[:len(chunk)], [src] * len(chunk), xs)):
            acc[p, 0 if sg > 0 else 1] = val
        print(f"finite differences: {min(j0 + 12, len(jobs))} of {len(jobs)} solves", flush=True)
    del sv
    _free()
    fd = (acc[:, 0] - acc[:, 1]) / (2 * dV)
    an = np.array([g[i] if kind == "V" else float(W[i] @ g) for (kind, i), _ in perturb])
    scale = np.abs(g).max()
    err = np.abs(fd - an)
    print("fd vs adjoint: max abs err", err.max(), "max |g|", scale, "max rel (shifts)",
          (err[48:] / np.abs(an[48:])).max(), flush=True)
    for ((kind, i), _), f, a in zip(perturb, fd, an):
        assert abs(f - a) <= 1e-6 * scale + 1e-5 * abs(a), (kind, i, f, a)
