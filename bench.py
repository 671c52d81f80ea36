"""Benchmark of the FDM energy-loop hot path on one MI355X.

One step = one group of G energy points of the Fe cluster (bcc Fe, 59 atoms,
R = 5.0 A, h = 0.2 A, 65,283 unknowns, 16 right-hand sides, complex128
system): assembly into the device factor storage, block-skyline UL
factorisation, substitutions (with FP64 refinement for the mixed solver) and
sigma(E).  Inputs (matrix values and sources of every energy) are resident in
HBM before the timed region.  Reports the mixed-precision solver's energy
points/s as `value`, with the FP64 drop-in's rate, the sigma(E) error of the
mixed solver against the FP64 one, the adjoint/forward time ratio, the
roofline of the trailing-update kernel and the CPU baseline (the oracle's
restatement of the reference Gaussian solver on a bounded sample).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--group G]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = ("energy points/s, mixed-prec vs FP64 GPU at ≤1e-6 rel σ(E) err; adjoint/fwd time")
PEAK = {"fp64": 78.6, "fp32": 157.3}  # MI355X dense MFMA TFLOP/s


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline(w, target_s: float, procs: int):
    """The oracle's restatement of mat_solve (mat_solve_gaussian.f90:67-358,
    single-threaded like the reference) on the same cluster: each process
    eliminates the last rows of a different energy point until about
    target_s of elimination work, and the rate is scaled by the fraction of
    the factorisation's multiply-adds those rows carry."""
    import multiprocessing as mp
    ctx = mp.get_context("spawn")
    with ctx.Pool(procs) as pool:
        res = pool.starmap(_cpu_worker, [(w.name, k, target_s) for k in range(procs)])
    rates = [r[0] for r in res]
    sample = res[0][1]
    return {"value": float(np.sum(rates)), "unit": "energy points/s", "cores": procs, "kind": "port",
            "sample": sample}


def _cpu_worker(wname, k, target_s):
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from fdmnes_amd import configs as cfg
    from oracle import oracle as O
    w = cfg.WORKLOADS[wname]
    cl = O.Cluster(w.R / cfg.BOHR, w.h / cfg.BOHR, w.lmaxso)
    lb1, lb2, _, _ = cl.profile()
    n = cl.nligne
    # per-row elimination work (complex multiply-adds), bottom-up
    lb1l = lb1.astype(np.int64)
    cum = np.concatenate([[0], np.cumsum(np.arange(1, n + 1) - lb1l)])  # sum_{j<=m} (j - lb1(j))
    ii = np.arange(1, n + 1)
    row_work = cum[lb2] - cum[ii]
    total = float(row_work.sum())
    # calibrate: about 0.5e9 complex MACs/s per core for this memory-bound loop
    tgt = target_s * 0.5e9
    acc, stop = 0.0, n
    while stop > 1 and acc < tgt:
        acc += row_work[stop - 1]
        stop -= 1
    stop += 1
    frac = float(row_work[stop - 1:].sum()) / total
    xyz = cl.xyz[: cl.npoint, :3]
    V = cfg.potential(w, xyz)
    E = float(w.energies_Ry[(k * 7) % w.n_energies])
    st, x, fr, tf, tt, nnz = cl.solve(V, E, w.eimag_Ry, cfg.ecinetic_out(E), True, True, False, inc_wave=False,
                                      row_stop=stop)
    t = tf + tt
    return frac / t, (f"rows {stop}..{n} (last {n - stop + 1} of {n}) of one energy point per process, "
                      f"{frac * 100:.3f}% of the factorisation's {total:.3e} complex MACs, {t:.1f} s each; "
                      f"rate scaled by that fraction")


def _cu_worker(k):
    """One whole energy point of the copper cluster (real arithmetic,
    Basereel, calcMatRow's incoming-wave right-hand sides) through the
    oracle's restatement of mat_solve, then its reduction to tau over the
    block of the nlmso innermost grid rows; also the same reduction of a
    partial-pivoting dense solve (LAPACK through numpy, untimed), the
    accuracy reference on this undamped system.  Returns (k, seconds,
    sigma_tau of the oracle, sigma_tau of the dense solve)."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from fdmnes_amd import configs as cfg
    from oracle import oracle as O
    w = cfg.CU_REAL
    cl = O.Cluster(w.R / cfg.BOHR, w.h / cfg.BOHR, w.lmaxso)
    V = cfg.potential(w, cl.xyz[: cl.npoint, :3])
    E = float(w.energies_Ry[k])
    t0 = time.time()
    st, x, fr, tf, tt, _ = cl.solve(V, E, 0.0, cfg.ecinetic_out(E), False, False, True)
    t = time.time() - t0
    if st:
        return k, t, float("nan"), float("nan")
    nso = cl.nligne - cl.nlmso
    taus, _, info = O.tau(x, nso, True, False, [(0, cl.nlmso)])
    A, b = cl.assemble_dense(V, E, 0.0, cfg.ecinetic_out(E), False, False, True)
    taud, _, _ = O.tau(np.linalg.solve(A, b), nso, True, False, [(0, cl.nlmso)])
    return k, t, float(-np.imag(np.trace(taus[0]))), float(-np.imag(np.trace(taud[0])))


def copper_cpu_start(procs: int):
    """Starts the copper cluster's CPU sample in the background: procs
    processes, one whole energy point each, spread over the 91 energies."""
    import multiprocessing as mp
    from fdmnes_amd import configs as cfg
    n = cfg.CU_REAL.n_energies
    ks = sorted({int(round(i * (n - 1) / max(procs - 1, 1))) for i in range(procs)})
    for v in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
        os.environ[v] = "1"  # one core per worker process (inherited at spawn)
    pool = mp.get_context("spawn").Pool(len(ks))
    return pool, pool.map_async(_cu_worker, ks)


def copper_gpu(cpu_job):
    """The copper cluster (fcc Cu, R = 3.0 A, h = 0.24 A, l <= 2, real
    arithmetic, Basereel, no broadening) at all 91 energies on the GPU:
    grouped FP64 solves of calcMatRow's incoming-wave right-hand sides (the
    reference's RHS, nlmso = 9), then tau in the real basis over the block of
    the nlmso innermost grid rows (invcomp of the outer-sphere
    normalisation, fdmgpu_tau_reduce); sigma_tau(E) = -Im sum_L tau_LL, the
    density of states the reference reads from Taull (mat.f90:393-396).
    With the CPU sample (the oracle, whole energy points) its rate and the
    per-energy difference of sigma_tau at the sampled energies."""
    import torch
    from fdmnes_amd import batch as B
    from fdmnes_amd import cluster as CL
    from fdmnes_amd import configs as cfg
    from fdmnes_amd.tau import tau_reduce
    w = cfg.CU_REAL
    cl = CL.Cluster(w.R / cfg.BOHR, w.h / cfg.BOHR, w.lmaxso)
    lb1, lb2 = cl.profile()
    V = cfg.potential(w, cl.points)
    E = w.energies_Ry
    vals, rhs = [], []
    for e in E:
        rp, col, v, b = cl.assemble(V, float(e), 0.0, cfg.ecinetic_out(float(e)), False, True)
        vals.append(torch.from_numpy(np.ascontiguousarray(v)).cuda())
        rhs.append(torch.from_numpy(np.ascontiguousarray(b)).cuda())
    G = 32
    sv = B.DeviceSolver(B.FP64, False, cl.nlmso)
    slots = [sv.add_system(lb1, lb2, rp, col) for _ in range(G)]
    xs = [torch.zeros_like(b) for b in rhs]
    nso = cl.nligne - cl.nlmso
    sig = np.zeros(len(E))

    def one_pass():
        for g0 in range(0, len(E), G):
            idx = list(range(g0, min(g0 + G, len(E))))
            sv.solve(slots[: len(idx)], [vals[i] for i in idx], [rhs[i] for i in idx], [xs[i] for i in idx])
            tau, _, _ = tau_reduce([xs[i] for i in idx], nso, True, False, [(0, cl.nlmso)])
            for i, t in zip(idx, tau):
                sig[i] = -float(torch.diagonal(t.view(cl.nlmso, cl.nlmso)).imag.sum())

    one_pass()  # captures the graphs
    torch.cuda.synchronize()
    t0 = time.time()
    one_pass()
    torch.cuda.synchronize()
    t_gpu = time.time() - t0
    res = {"workload": w.name, "nligne": cl.nligne, "nrhs": cl.nlmso, "energies": len(E),
           "gpu_fp64_points_per_s": len(E) / t_gpu, "sigma_tau": [float(v) for v in sig],
           "energies_eV": [float(e * cfg.RYDB) for e in E]}
    log(f"copper: {len(E)} energies on the GPU in {t_gpu:.2f}s ({len(E) / t_gpu:.1f} points/s)")
    del sv, xs, vals, rhs
    torch.cuda.empty_cache()
    if cpu_job is not None:
        pool, job = cpu_job
        out = job.get()
        pool.close()
        pool.join()
        ts = [t for _, t, _, _ in out]
        rel = [abs(sig[k] - sc) / abs(sc) for k, _, sc, _ in out]
        rel_d = [abs(sig[k] - sd) / abs(sd) for k, _, _, sd in out]
        rel_od = [abs(sc - sd) / abs(sd) for k, _, sc, sd in out]
        res["cpu"] = {"kind": "port", "cores": len(out), "points": len(out),
                      "mean_s_per_point": float(np.mean(ts)), "points_per_s": len(out) / float(np.mean(ts)),
                      "sample": f"{len(out)} of the {len(E)} energies, one whole point (assembly and mat_solve "
                                f"restatement) per process, concurrent",
                      "energies_eV": [float(E[k] * cfg.RYDB) for k, _, _, _ in out],
                      "sigma_tau_rel_diff_gpu_vs_oracle": [float(r) for r in rel],
                      "sigma_tau_rel_diff_gpu_vs_pivoted_dense": [float(r) for r in rel_d],
                      "sigma_tau_rel_diff_oracle_vs_pivoted_dense": [float(r) for r in rel_od]}
        res["gpu_over_cpu"] = res["gpu_fp64_points_per_s"] / res["cpu"]["points_per_s"]
        log(f"copper CPU: {len(out)} points, {np.mean(ts):.1f} s each; sigma_tau max rel diff: GPU vs oracle "
            f"{max(rel):.2e}, GPU vs pivoted dense {max(rel_d):.2e}, oracle vs pivoted dense {max(rel_od):.2e}")
    return res


def batch_waves(budget_gb: float):
    """The three iron clusters of uneven size (R = 3, 4, 5 A, h = 0.2 A;
    86 + 85 + 85 = 256 energy points, FP64, 16 right-hand sides each) in
    HBM-sized waves of grouped solves: the solver holds n_c slots of cluster
    c, chosen so the factor storage fits the budget, at most 32 slots, with
    the fewest waves; each wave is one grouped graph-captured solve of the
    next n_c points of every cluster.  Host assembly and upload of all
    points precede the timed region."""
    import itertools as it
    import torch
    from fdmnes_amd import batch as B
    from fdmnes_amd import cluster as CL
    from fdmnes_amd import configs as cfg
    cls, pts = [], []
    for ci, w in enumerate(cfg.BATCH):
        cl = CL.Cluster(w.R / cfg.BOHR, w.h / cfg.BOHR, w.lmaxso)
        lb1, lb2 = cl.profile()
        V = cfg.potential(w, cl.points)
        s = torch.from_numpy(cfg.sources(w, cl.points, cl.nligne).astype(np.complex128)).cuda()
        vals = []
        for e in w.energies_Ry:
            rp, col, v, _ = cl.assemble(V, float(e), w.eimag_Ry, cfg.ecinetic_out(float(e)), True, False)
            vals.append(torch.from_numpy(v.astype(np.complex128)).cuda())
        cls.append(dict(w=w, cl=cl, lb1=lb1, lb2=lb2, rp=rp, col=col, s=s, vals=vals))
        pts.append(len(vals))
    # factor bytes per slot of each cluster
    nrhs = cls[0]["s"].shape[1]
    per = []
    for c in cls:
        sv = B.DeviceSolver(B.FP64, True, nrhs)
        sv.add_system(c["lb1"], c["lb2"], c["rp"], c["col"])
        per.append(sv.stats()["factor_bytes"] / 1e9)
        del sv
    torch.cuda.empty_cache()
    best = None
    for n in it.product(*[range(1, 33)] * len(cls)):
        if sum(n) > 32 or sum(a * b for a, b in zip(n, per)) > budget_gb:
            continue
        waves = max(-(-p // k) for p, k in zip(pts, n))
        key = (waves, sum(a * b for a, b in zip(n, per)))
        if best is None or key < best[0]:
            best = (key, n)
    (waves, gb), nslot = best
    sv = B.DeviceSolver(B.FP64, True, nrhs)
    slots = [[sv.add_system(c["lb1"], c["lb2"], c["rp"], c["col"]) for _ in range(k)] for c, k in zip(cls, nslot)]
    xs = [[torch.zeros_like(c["s"]) for _ in c["vals"]] for c in cls]
    sig = [np.zeros(p) for p in pts]

    def wave(v):
        sl, va, b, x, where = [], [], [], [], []
        for ci, c in enumerate(cls):
            for j in range(nslot[ci]):
                k = v * nslot[ci] + j
                if k < pts[ci]:
                    sl.append(slots[ci][j])
                    va.append(c["vals"][k])
                    b.append(c["s"])
                    x.append(xs[ci][k])
                    where.append((ci, k))
        sv.solve(sl, va, b, x)
        for (ci, k), val in zip(where, sv.sigma(sl, b, x)):
            sig[ci][k] = val

    wave(0)  # warm-up: captures the first wave's graphs (later masks are captured in the timed run)
    torch.cuda.synchronize()
    t0 = time.time()
    for v in range(waves):
        wave(v)
    torch.cuda.synchronize()
    t = time.time() - t0
    res = {"points": int(sum(pts)), "points_per_cluster": pts, "nligne": [c["cl"].nligne for c in cls],
           "slots_per_cluster": list(nslot), "waves": waves, "factor_GB_in_flight": gb,
           "factor_GB_per_slot": per, "seconds": t, "points_per_s": sum(pts) / t,
           "workloads": [c["w"].name for c in cls]}
    log(f"batch: {sum(pts)} points of {len(cls)} clusters in {waves} waves (slots {nslot}, {gb:.0f} GB), "
        f"{t:.1f}s -> {sum(pts) / t:.2f} points/s")
    del sv, xs, cls
    torch.cuda.empty_cache()
    return res


def adjoint_recompute(w, n_energies: int, slots: int):
    """Adjoint of sigma(E) over n_energies energies of the iron cluster
    through the torch.autograd.Function (fdmnes_amd.adjoint): forward solves
    in groups of `slots` (FP64, the last group's factors stay resident, the
    checkpoint), then the backward pass: transposed substitutions on the
    resident factors, refactorisation of every other point (recompute)
    before its transposed substitutions.  Host assembly is inside both
    passes.  Reports backward / forward time, and from the per-point costs
    the same ratio for 200 energies with the same checkpoint."""
    import torch
    from fdmnes_amd import adjoint as AD
    from fdmnes_amd import cluster as CL
    from fdmnes_amd import configs as cfg
    cl = CL.Cluster(w.R / cfg.BOHR, w.h / cfg.BOHR, w.lmaxso)
    s = cfg.sources(w, cl.points, cl.nligne)
    eng = AD.SigmaEngine(cl, s, w.eimag_Ry, slots=slots)
    E = w.energies_Ry[np.linspace(0, len(w.energies_Ry) - 1, n_energies).round().astype(int)]
    V = torch.tensor(cfg.potential(w, cl.points), dtype=torch.float64, requires_grad=True)
    eng.forward(V.detach().numpy(), E[:slots])  # warm-up: graphs
    torch.cuda.synchronize()
    t0 = time.time()
    sig = AD.SigmaFunction.apply(V, eng, E)
    torch.cuda.synchronize()
    t_f = time.time() - t0
    eng.recomputed = 0
    t0 = time.time()
    sig.sum().backward()
    torch.cuda.synchronize()
    t_b = time.time() - t0
    nrec = eng.recomputed
    nres = n_energies - nrec
    t_fp = t_f / n_energies
    # 200 energies with the same checkpoint: 200 - slots points recomputed;
    # the whole backward time is charged to the recomputed points (an upper
    # bound: the resident points' transposed solves are in it too)
    proj = (max(200 - slots, 0) * t_b / max(nrec, 1)) / (200 * t_fp)
    res = {"energies": n_energies, "checkpoint_slots": slots, "recomputed": nrec, "resident": nres,
           "forward_s": t_f, "backward_s": t_b, "adjoint_over_forward": t_b / t_f,
           "projected_adjoint_over_forward_200": proj, "grad_norm": float(V.grad.norm())}
    log(f"adjoint: {n_energies} energies, {slots} checkpointed, {nrec} recomputed: forward {t_f:.1f}s, "
        f"backward {t_b:.1f}s, ratio {t_b / t_f:.3f} (200 energies: {proj:.3f})")
    del eng, V, sig
    torch.cuda.empty_cache()
    return res


def dropin_rates(cl, lb1, lb2, V, s, w):
    """Seconds per energy point through fdmgpu_mat_solve / fdmgpu_mat_solve_mixed
    (host buffers in and out, device set-up per call), the Fortran drop-in's
    path; the second of two calls is reported."""
    from fdmnes_amd import configs as cfg
    from fdmnes_amd import solver as SV
    E = float(w.energies_Ry[len(w.energies_Ry) // 2])
    rp, col, vals, _ = cl.assemble(V, E, w.eimag_Ry, cfg.ecinetic_out(E), True, False)
    res = {}
    for tag, fn in (("fp64", SV.mat_solve), ("mixed", SV.mat_solve_mixed)):
        for _ in range(2):
            t0 = time.time()
            fn(lb1, lb2, rp, col, vals, s, True)
            dt = time.time() - t0
        res[f"dropin_{tag}_s_per_point"] = dt
        log(f"drop-in {tag}: {dt:.2f} s per point (host buffers)")
    return res


def assembly_rates(cl, V, w, G):
    """Per-energy row assembly: the device assembler (fdmgpu_assembler_run,
    G points per call, HIP events; fresh buffers, and buffers updated in
    place) beside the host assembler (fdmgpu_cluster_assemble, one core) on
    the same points.  Algorithmic HBM
    bytes per point: every value written (16 B) and its base entry read
    (8 B), the diagonal positions and V read (16 B per grid point), the
    outer-sphere tables read once per call."""
    import torch
    from fdmnes_amd import configs as cfg
    from fdmnes_amd.assemble import DeviceAssembler
    t = cl.assembly_tables()
    E = [float(w.energies_Ry[(k * 13) % len(w.energies_Ry)]) for k in range(G)]
    ens = [(e, w.eimag_Ry, cfg.ecinetic_out(e)) for e in E]
    stream = torch.cuda.current_stream()
    a = DeviceAssembler(cl, max_points=min(G, 16))
    Vd = torch.from_numpy(V).cuda()
    vals = [torch.empty(cl.nnz, dtype=torch.complex128, device="cuda") for _ in range(G)]
    a.run([Vd] * G, ens, True, False, vals)
    reps = 20

    def timed(update_only):
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            a.run([Vd] * G, ens, True, False, vals, update_only=update_only)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3 / (reps * G)

    dev_s, upd_s = timed(False), timed(True)
    h0 = time.time()
    nh = 3
    for k in range(nh):
        cl.assemble(V, *ens[k], True, False)
    host_s = (time.time() - h0) / nh
    nbytes = cl.nnz * 24 + cl.npoint * 16 + (t["nrow_out"] * 12 + t["nlink"] * 12 + t["nb"] * cl.nlmso * 8) / G
    res = {"device_us_per_point": dev_s * 1e6, "device_update_us_per_point": upd_s * 1e6,
           "host_ms_per_point": host_s * 1e3, "host_cores": 1,
           "bytes_per_point": nbytes, "achieved_GBps": nbytes / dev_s / 1e9,
           "frac_hbm": nbytes / dev_s / 8e12, "points_per_call": G}
    log(f"assembly: device {dev_s * 1e6:.1f} us/point ({upd_s * 1e6:.1f} updating), host {host_s * 1e3:.1f} ms/point")
    return res


def tau_rates(xs, cl, G):
    """Post-solve reduction of the last group's solutions to tau
    (fdmgpu_tau_reduce, HIP events): the workload's complex basis (the
    contraction only), and the real-basis path (invcomp of the outer-sphere
    block, then the contraction) on the same data.  One atom block of nlmso
    rows; only tau (nlmso^2 complex entries per point) would leave the GPU
    instead of the solution."""
    import torch
    from fdmnes_amd.tau import tau_reduce
    atoms = [(0, cl.nlmso)]
    nso = cl.nligne - cl.nlmso
    stream = torch.cuda.current_stream()
    reps = 20
    res = {"points_per_call": G, "atom_rows": cl.nlmso}
    for tag, basereel in (("complex_basis", False), ("real_basis", True)):
        tau_reduce(xs, nso, basereel, False, atoms, check=False)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            tau_reduce(xs, nso, basereel, False, atoms, check=False)
        e1.record(stream)
        torch.cuda.synchronize()
        res[f"{tag}_us_per_point"] = e0.elapsed_time(e1) * 1e3 / (reps * G)
    res["bytes_out_per_point"] = cl.nlmso * cl.nlmso * 16
    res["solution_bytes_per_point"] = cl.nligne * cl.nlmso * 16
    log(f"tau: {res['complex_basis_us_per_point']:.1f} us/point (complex basis), "
        f"{res['real_basis_us_per_point']:.1f} us/point (real basis with invcomp)")
    return res


def mmm_rates(w, G):
    """Mirror-symmetric potential (point group mmm): the whole cluster's
    mixed-precision grouped solve of G energies against the same energies
    as the eight representations' systems on the wedge x, y, z >= 0
    (fdmgpu_cluster_create_mmm), 32 systems per grouped solve (four
    energies of all eight), both with calcMatRow's incoming-wave RHS, HIP
    events around the timed solves (one warm-up group each).  Also the
    largest difference between the two, u / rvol against the whole
    solution, over every wedge row and harmonic of the first energy."""
    import torch
    from fdmnes_amd import _lib
    from fdmnes_amd import batch as B
    from fdmnes_amd import cluster as CL
    from fdmnes_amd import configs as cfg
    R, h = w.R / cfg.BOHR, w.h / cfg.BOHR
    full = CL.Cluster(R, h, w.lmaxso)
    E = [float(w.energies_Ry[(k * 13) % len(w.energies_Ry)]) for k in range(G)]
    ens = [(e, w.eimag_Ry, cfg.ecinetic_out(e)) for e in E]
    stream = torch.cuda.current_stream()

    def timed(fn):
        fn()  # warm-up: builds the graphs
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3

    V = cfg.potential(w, full.points, mirror_symmetric=True)
    lb1, lb2 = full.profile()
    sv = B.DeviceSolver(B.FP32, True, full.nlmso)
    slots, vals, rhs = [], [], []
    for e in ens:
        rp, col, v, b = full.assemble(V, *e, True, False)
        slots.append(sv.add_system(lb1, lb2, rp, col))
        vals.append(torch.from_numpy(v).cuda())
        rhs.append(torch.from_numpy(b).cuda())
    from fdmnes_amd import autotune
    xs = [torch.zeros_like(b) for b in rhs]
    autotune.tune(sv, slots, vals, rhs, xs, lb1, lb2, B.FP32, log=log)
    t_full = timed(lambda: sv.solve(slots, vals, rhs, xs))
    x0 = xs[0].cpu().numpy()
    del sv, vals, rhs, xs
    torch.cuda.empty_cache()
    reps = []
    for g in range(8):
        try:
            reps.append(CL.Cluster(R, h, w.lmaxso, irrep=g))
        except _lib.FdmGpuError:
            pass
    nrhs = max(c.nlmso for c in reps)
    per = 32 // len(reps)  # energies per grouped solve of up to 32 systems
    sv = B.DeviceSolver(B.FP32, True, nrhs)
    prof = [c.profile() for c in reps]
    pots = [cfg.potential(w, c.points, mirror_symmetric=True) for c in reps]
    groups = []
    rslots = None
    for e0 in range(0, G, per):
        vals, rhs = [], []
        for c, (l1, l2), Vc in zip(reps, prof, pots):
            for e in ens[e0:e0 + per]:
                rp, col, v, b = c.assemble(Vc, *e, True, False)
                bp = np.zeros((c.nligne, nrhs), np.complex128)
                bp[:, :c.nlmso] = b
                vals.append(torch.from_numpy(v).cuda())
                rhs.append(torch.from_numpy(bp).cuda())
        if rslots is None:  # slot ci * per + j: representation ci, j-th energy of a group
            rslots = [sv.add_system(l1, l2, *c.assemble(Vc, *ens[0], True, False)[:2])
                      for c, (l1, l2), Vc in zip(reps, prof, pots) for _ in range(per)]
        k = len(ens[e0:e0 + per])
        gslots = [rslots[ci * per + j] for ci in range(len(reps)) for j in range(k)]
        groups.append((gslots, vals, rhs, [torch.zeros_like(b) for b in rhs]))

    # schedule for the reduced group, tuned (or taken from the cache) on the
    # first group, keyed by the first representation's profile
    gs0, v0, r0, x0r = groups[0]
    ent = autotune.tune(sv, gs0, v0, r0, x0r, prof[0][0], prof[0][1], B.FP32, log=log)

    def reduced():
        for gslots, vals, rhs, xs in groups:
            sv.solve(gslots, vals, rhs, xs)

    t_red = timed(reduced)
    row = {tuple(k): i for i, k in enumerate(np.round(full.points / h).astype(int))}
    nsof = full.nligne - full.nlmso
    dmax = 0.0
    for ci, c in enumerate(reps):
        u = groups[0][3][ci * min(per, G)].cpu().numpy()
        hm = c.harmonics
        rows = [row[tuple(k)] for k in np.round(c.points / h).astype(int)]
        dmax = max(dmax, np.abs(u[:c.npoint, :c.nlmso] / c.rvol[:, None] - x0[rows][:, hm]).max(),
                   np.abs(u[c.npoint:, :c.nlmso] - x0[nsof + hm][:, hm]).max())
    res = {"points": G, "whole_points_per_s": G / t_full, "reduced_points_per_s": G / t_red,
           "speedup": t_full / t_red, "max_rel_diff": dmax / np.abs(x0).max(),
           "representations": [int(c.irrep) for c in reps], "rows_per_representation": [c.nligne for c in reps],
           "whole_rows": full.nligne, "reduced_schedule": ent["best"]}
    log(f"mmm: whole {G / t_full:.3f} points/s, eight representations {G / t_red:.3f} points/s, "
        f"max rel diff {res['max_rel_diff']:.2e}")
    return res


def atoms_rates(w, G):
    """The energy loop with muffin-tin atoms on the benchmark's grid: the
    Fe sites within 3 A of the absorber as atoms (rmt 1.0 A, lmax 2,
    screened-Coulomb spherical potentials), per group of G energies the
    device radial solver (regular solution, phiato, the absorber's radial
    integrals), device assembly, the mixed-precision grouped solve, tau
    and the absorber's E1E1 / E1E2 / E2E2 tensors (K edge).  One warm-up
    group, then two timed groups (wall clock around synchronised runs);
    the radial solver and the tensor kernel also timed alone (HIP events,
    20 calls of G points)."""
    import torch
    from fdmnes_amd import cluster as CL
    from fdmnes_amd import configs as cfg
    from fdmnes_amd.energy_loop import FP32, EnergyLoop
    from fdmnes_amd.sphere import DeviceSphere, atoms_of
    from fdmnes_amd.tensor import DeviceTensor, no_symmetry_tables
    sites = cfg.lattice_sites(w.lattice, w.a, 3.0) / cfg.BOHR
    rmt = np.full(len(sites), 1.0 / cfg.BOHR)
    lmax = np.full(len(sites), 2, np.int32)
    t0 = time.perf_counter()
    cl = CL.Cluster(w.R / cfg.BOHR, w.h / cfg.BOHR, w.lmaxso, atoms=(sites, rmt, lmax))
    t_build = time.perf_counter() - t0
    lam = w.screening / cfg.BOHR
    grids, pots = [], []
    for a in range(len(sites)):
        rr = [1e-4]
        while rr[-1] < 1.3 * rmt[a] + 2 * w.h / cfg.BOHR:
            rr.append(rr[-1] * 1.03)
        rr = np.array(rr)
        grids.append(rr)
        pots.append(-2.0 * w.Z * np.exp(-rr / lam) / rr - 0.5)
    sph = DeviceSphere(atoms_of(cl, grids, pots, [w.Z] * len(sites), rmt, sites))
    psii = (grids[0] * np.exp(-w.Z * grids[0]))[None]
    sph.set_integrals(0, psii, [7112.0 / cfg.RYDB], 1.1 * rmt[0])
    nl0 = int((lmax[0] + 1) ** 2)
    ten = DeviceTensor(1, 0, 1, 2, no_symmetry_tables(), sph.rof_layout, 0, nl0, numat=w.Z)
    V = cfg.potential(w, cl.points)
    E = [float(w.energies_Ry[(k * 7) % len(w.energies_Ry)]) for k in range(3 * G)]
    ens = [(e, w.eimag_Ry, cfg.ecinetic_out(e)) for e in E]
    loop = EnergyLoop(cl, precision=FP32, group=G, sphere=sph, tensor=ten)
    loop.run(V, ens[:G], energ=E[:G])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = loop.run(V, ens[G:], energ=E[G:])
    torch.cuda.synchronize()
    t_loop = time.perf_counter() - t0
    sens = [(e[2], e[1], e[0]) for e in ens[:G]]
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(sph.stream)
    for _ in range(20):
        phi, rof = sph.run(sens, energ=E[:G])
    ev1.record(sph.stream)
    torch.cuda.synchronize()
    t_sph = ev0.elapsed_time(ev1) / 1e3 / 20
    tau = [torch.zeros(sum(int(cl.atom(a)["nlmsa"]) ** 2 for a in range(cl.natome)), dtype=torch.complex128,
                       device="cuda") + 0.1j for _ in range(G)]
    ev0.record(ten.stream)
    for _ in range(20):
        ten.run(rof, tau)
    ev1.record(ten.stream)
    torch.cuda.synchronize()
    t_ten = ev0.elapsed_time(ev1) / 1e3 / 20
    log(f"atoms: {cl.natome} atoms, {cl.nligne} rows, {2 * G} points in {t_loop:.2f} s")
    return {"natome": int(cl.natome), "nligne": int(cl.nligne), "nlmso": int(cl.nlmso), "build_s": t_build,
            "points": 2 * G, "points_per_s": 2 * G / t_loop, "sphere_us_per_point": 1e6 * t_sph / G,
            "tensor_us_per_point": 1e6 * t_ten / G, "sweeps": [int(v) for v in out["sweeps"]],
            "secdd_trace_first": float(np.trace(out["secdd"][0][0]).real)}


def traffic_from_profile():
    """HBM bytes per k_trail launch measured by rocprofv3 PMC runs (FETCH_SIZE,
    WRITE_SIZE) of this benchmark, scaled by the same counters on a plain
    copy of known size; committed under profiles/.  None if absent."""
    p = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic_k_trail_fe59_R5.0_mixed.json")
    try:
        with open(p) as f:
            return json.load(f)["bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        return None


def roofline(sv32, slots32, vals, src, G, peak_key="fp32", kernel="k_trail<float,complex>"):
    import torch
    sv32.options(use_graphs=False, time_updates=True)
    grp = list(range(G))
    xs = [torch.zeros_like(src) for _ in range(G)]
    sv32.solve(slots32, [vals[i] for i in grp], [src] * G, xs)
    st = sv32.stats()
    upd_s = st["update_ms"] / 1e3
    # algorithmic flops of the update launches: the element-profile multiply-adds
    # of the G factorisations times the trailing updates' share of the tile-level
    # work (the rest is panels and triangular block products)
    share = st["tile_macs"] / (st["tile_macs"] + st["tile_other_macs"])
    alg_flops = 8.0 * st["profile_macs"] * share
    ach = alg_flops / upd_s / 1e12
    roof = {"bound": "mfma", "achieved": ach, "peak": PEAK[peak_key], "unit": "TFLOP/s", "frac": ach / PEAK[peak_key],
            "traffic": traffic_from_profile() if peak_key == "fp32" else None,
            "kernel": f"{kernel} (trailing update: cross + group launches)",
            "launches": st["update_launches"], "avg_launch_us": 1e6 * upd_s / max(st["update_launches"], 1),
            "algorithmic_flops_per_launch": alg_flops / max(st["update_launches"], 1),
            "update_share_of_factorisation": share,
            "algorithmic_bytes_per_launch": st["update_bytes"] / max(st["update_launches"], 1),
            "executed_tile_flops": 8.0 * st["tile_macs"]}
    return roof


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--group", type=int, default=12, help="energy points per step (grouped launch)")
    ap.add_argument("--workload", default="fe59_R5.0_h0.2_c128")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-procs", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--no-roofline", action="store_true", help="skip the per-launch timed pass (profiler runs)")
    ap.add_argument("--no-assembly", action="store_true", help="skip the per-energy assembly timing")
    ap.add_argument("--no-tau", action="store_true", help="skip the post-solve tau reduction timing")
    ap.add_argument("--no-graphs", action="store_true", help="launch eagerly instead of replaying HIP graphs")
    ap.add_argument("--overlap", type=int, default=None, choices=(0, 1),
                    help="per-step overlap of the trailing update (fdmgpu_solver_overlap); default: the library's")
    ap.add_argument("--atoms", action="store_true",
                    help="also time the energy loop with muffin-tin atoms, radial solver and tensors")
    ap.add_argument("--mmm", action="store_true",
                    help="also time the mirror-symmetric workload: whole cluster against its eight representations")
    ap.add_argument("--dropin", action="store_true",
                    help="also time the host-buffer drop-in entry points (PCIe and set-up included), one point each")
    ap.add_argument("--fp64-steps", type=int, default=4,
                    help="timed groups of the FP64 comparison leg (at most --steps), after one warm-up group")
    ap.add_argument("--adjoint-energies", type=int, default=24,
                    help="energies of the adjoint leg (forward, then backward with checkpoint/recompute)")
    ap.add_argument("--adjoint-slots", type=int, default=12, help="checkpointed factor slots of the adjoint leg")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the three-cluster 256-point wave leg")
    ap.add_argument("--batch-budget-gb", type=float, default=240.0)
    ap.add_argument("--no-copper", action="store_true", help="skip the copper cluster leg (GPU and CPU sample)")
    ap.add_argument("--copper-cpu-procs", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs shard energy points "
                                   "across processes and are not part of it"}))
        sys.exit(2)

    import torch

    from fdmnes_amd import batch as B
    from fdmnes_amd import cluster as CL
    from fdmnes_amd import configs as cfg

    w = cfg.WORKLOADS[args.workload]
    G, K, W = args.group, args.steps, args.warmup
    t0 = time.time()
    cl = CL.Cluster(w.R / cfg.BOHR, w.h / cfg.BOHR, w.lmaxso)
    lb1, lb2 = cl.profile()
    pts = cl.points
    V = cfg.potential(w, pts)
    s = cfg.sources(w, pts, cl.nligne).astype(np.complex128)
    log(f"cluster {w.name}: nligne={cl.nligne} nnz={cl.nnz} built in {time.time() - t0:.1f}s")
    energies = w.energies_Ry
    n_pts = (W + K) * G
    idx = [(k * 13) % len(energies) for k in range(n_pts)]
    dev = torch.device("cuda")
    vals, rp, col = [], None, None
    for k in idx:
        E = float(energies[k])
        rp, col, v, _ = cl.assemble(V, E, w.eimag_Ry, cfg.ecinetic_out(E), True, False)
        vals.append(torch.from_numpy(v.astype(np.complex128)).to(dev))
    src = torch.from_numpy(s).to(dev)
    log(f"assembled and uploaded {n_pts} energy points in {time.time() - t0:.1f}s")

    def run(precision, tag, K, W):
        sv = B.DeviceSolver(precision, True, s.shape[1])
        slots = [sv.add_system(lb1, lb2, rp, col) for _ in range(G)]
        xs = [torch.zeros_like(src) for _ in range(G)]
        if not args.no_tune:  # schedule from the autotuner's cache, tuned on the first group if absent
            from fdmnes_amd import autotune
            ent = autotune.tune(sv, slots, [vals[i] for i in range(G)], [src] * G, xs, lb1, lb2, precision,
                                log=log)
            tuned[tag] = ent["best"]
        if args.overlap is not None:
            sv.overlap(args.overlap)
        if args.no_graphs:
            sv.options(use_graphs=False)
        sig = np.zeros(n_pts)
        sweeps = []
        stream = torch.cuda.current_stream()
        for step in range(W + K):
            if step == W:
                torch.cuda.synchronize()
                e0 = torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                h0 = time.time()
            grp = list(range(step * G, (step + 1) * G))
            fr, sw, rel = sv.solve(slots, [vals[i] for i in grp], [src] * G, xs)
            sweeps += sw
            sig[grp] = sv.sigma(slots, [src] * G, xs)
        e1 = torch.cuda.Event(enable_timing=True)
        e1.record(stream)
        torch.cuda.synchronize()
        wall = time.time() - h0
        gpu_ms = e0.elapsed_time(e1)
        log(f"{tag}: {K} steps x {G} points in {wall:.2f}s (events {gpu_ms / 1e3:.2f}s)")
        return sv, slots, xs, sig, wall, sweeps

    tuned = {}
    # mixed precision (headline)
    sv32, slots32, _, sig32, wall32, sweeps32 = run(B.FP32, "mixed", K, W)
    st32 = sv32.stats()
    rate32 = K * G / wall32
    # roofline of the dominant kernel: the FP32 trailing update, timed per
    # launch with HIP events on the solver's stream (one instrumented group)
    roof = None if args.no_roofline else roofline(sv32, slots32, vals, src, G)
    del sv32  # its factor storage makes room for the FP64 solver's
    torch.cuda.empty_cache()
    out = {}
    # the copper cluster's CPU sample runs on the host cores beside the FP64 legs
    cu_job = None
    if not args.no_copper and not args.no_cpu:
        cu_procs = args.copper_cpu_procs or max(1, min(15, (os.cpu_count() or 2) - 1))
        log(f"copper CPU sample on {cu_procs} processes (background)")
        cu_job = copper_cpu_start(cu_procs)
    if not args.no_fp64:
        K64, W64 = min(K, args.fp64_steps), min(W, 1)
        sv64, slots64, xs64, sig64, wall64, _ = run(B.FP64, "fp64", K64, W64)
        rate64 = K64 * G / wall64
        n64 = (W64 + K64) * G  # the FP64 leg solved the first n64 of the mixed leg's points
        err = np.abs(sig32[:n64] - sig64[:n64]) / np.maximum(np.abs(sig64[:n64]), 1e-300)
        out["sigma_rel_err_per_energy"] = {"reference": "FP64 GPU solver (pinned to the oracle at <= 1e-10)",
                                           "energies_eV": [float(energies[k] * cfg.RYDB) for k in idx[:n64]],
                                           "rel_err": [float(e) for e in err]}
        out["fp64_steps"], out["fp64_warmup"] = K64, W64
        # adjoint: transposed solves with the stored FP64 factors of the last group
        grads = [torch.zeros(cl.nligne, dtype=torch.float64, device=dev) for _ in range(G)]
        torch.cuda.synchronize()
        a0 = time.time()
        sv64.adjoint(slots64, [src] * G, xs64, grads)
        torch.cuda.synchronize()
        t_adj = time.time() - a0
        t_fwd = wall64 / K64
        if not args.no_roofline:
            r64 = roofline(sv64, slots64, vals, src, G, "fp64", "k_trail<double,complex>")
            out["fp64_roofline"] = {k2: r64[k2] for k2 in ("achieved", "peak", "unit", "frac", "avg_launch_us")}
        out.update({"fp64_points_per_s": rate64, "mixed_over_fp64": rate32 / rate64,
                    "max_rel_sigma_err": float(err.max()), "adjoint_resident_over_forward": t_adj / t_fwd,
                    "adjoint_s_per_group": t_adj, "fp64_ms_per_step": 1e3 * t_fwd})
        if not args.no_tau:
            out["tau"] = tau_rates(xs64, cl, G)
        del sv64, xs64, grads
        torch.cuda.empty_cache()
    if not args.no_adjoint:
        out["adjoint"] = adjoint_recompute(w, args.adjoint_energies, args.adjoint_slots)
        out["adjoint_over_forward"] = out["adjoint"]["adjoint_over_forward"]
    if not args.no_batch:
        out["batch"] = batch_waves(args.batch_budget_gb)
    if not args.no_copper:
        out["copper"] = copper_gpu(cu_job)
    if not args.no_assembly:
        out["assembly"] = assembly_rates(cl, V, w, G)
    if args.dropin:
        out.update(dropin_rates(cl, lb1, lb2, V, s, w))
    if args.mmm:
        out["mmm"] = mmm_rates(w, G)
    if args.atoms:
        out["atoms"] = atoms_rates(w, min(G, 6))
    cpu = None
    if not args.no_cpu:
        procs = args.cpu_procs or min(16, os.cpu_count() or 1)
        log(f"cpu baseline on {procs} processes")
        cpu = cpu_baseline(w, args.cpu_seconds, procs)
    res = {
        "metric": METRIC, "value": rate32, "unit": "energy points/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": 1e3 * wall32 / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 factors + c128 refinement", "data": "synthetic",
        "config": {"workload": w.name, "nligne": cl.nligne, "nrhs": int(s.shape[1]), "points_per_step": G,
                   "energies_eV": [float(e * cfg.RYDB) for e in energies[idx[W * G:]]][:8],
                   "factor_bytes_per_point": st32["factor_bytes"] / G, "profile_cmacs_per_point": st32["profile_macs"] / G},
        "refine_sweeps": sweeps32[W * G:], "schedule": tuned,
        "roofline": roof, "cpu_baseline": cpu,
    }
    res.update(out)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
