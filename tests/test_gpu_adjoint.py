"""d sigma(E)/dV through the torch.autograd.Function harness: against central
finite differences (dV = 1e-6 Ry) on random parameters, against an FP64 CPU
dense-solve gradient, per-atom shift gradients by the chain rule, and the
checkpoint/recompute path (more energies than resident factor slots)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
BOHR = 0.52917721067


def _engine(slots):
    from fdmnes_amd import adjoint as AD
    from fdmnes_amd import cluster as CL
    from fdmnes_amd import configs as cfg
    w = cfg.FE_FP64
    cl = CL.Cluster(1.5 / BOHR, 0.25 / BOHR, 2)
    pts = cl.points
    s = cfg.sources(w, pts, cl.nligne)[:, :9]
    return AD.SigmaEngine(cl, s, w.eimag_Ry, slots=slots), cl, pts, s, w


def test_autograd_gradient_fd_dense_and_recompute():
    import torch
    from fdmnes_amd import adjoint as AD
    from fdmnes_amd import configs as cfg
    from tests.helpers import dense_matrix
    eng, cl, pts, s, w = _engine(slots=2)
    V0 = cfg.potential(w, pts)
    energies = np.array([-0.3, 0.2, 0.9, 1.7, 2.6])  # 5 points, 2 slots: recompute needed
    V = torch.tensor(V0, dtype=torch.float64, requires_grad=True)
    sig = AD.SigmaFunction.apply(V, eng, energies)
    wts = torch.tensor([1.0, -0.5, 0.25, 2.0, 0.1], dtype=torch.float64)
    (sig * wts).sum().backward()
    g = V.grad.numpy()
    assert eng.recomputed > 0
    print("recomputed", eng.recomputed, flush=True)
    # FP64 CPU dense gradient
    g_ref = np.zeros_like(V0)
    for E, wt in zip(energies, wts.numpy()):
        rp, col, vals, _ = cl.assemble(V0, float(E), w.eimag_Ry, cfg.ecinetic_out(float(E)), True, False)
        lb1, lb2 = cl.profile()
        A = dense_matrix(dict(lb1=lb1, rp=rp, col=col, vals=vals))
        x = np.linalg.solve(A, s)
        lam = np.linalg.solve(A.T, s)
        g_ref += wt * np.imag(np.sum(lam * x, axis=1))[: cl.npoint]
    assert np.abs(g - g_ref).max() <= 1e-9 * np.abs(g_ref).max()
    # central finite differences on random grid points
    rng = np.random.default_rng(5)
    dV = 1e-6
    for i in rng.choice(cl.npoint, 6, replace=False):
        vals = []
        for sg in (1, -1):
            Vp = V0.copy()
            Vp[i] += sg * dV
            sp, _ = eng.forward(Vp, energies)
            vals.append(float(np.dot(sp, wts.numpy())))
        fd = (vals[0] - vals[1]) / (2 * dV)
        assert abs(fd - g[i]) <= 1e-6 * np.abs(g).max(), (i, fd, g[i])
    # per-atom shift parameters: d/d shift_a = sum_i w_a(i) d/dV_i
    W = cfg.atom_weights(w, pts)
    ga = W @ g
    a = int(np.argmax(np.abs(ga)))
    vals = []
    for sg in (1, -1):
        sh = np.zeros(W.shape[0])
        sh[a] = sg * dV
        sp, _ = eng.forward(cfg.potential(w, pts, shifts=sh), energies)
        vals.append(float(np.dot(sp, wts.numpy())))
    fd = (vals[0] - vals[1]) / (2 * dV)
    assert abs(fd - ga[a]) <= 1e-6 * abs(ga[a]) + 1e-9, (fd, ga[a])
