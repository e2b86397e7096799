"""Benchmark of the MI355X SIE solve loop.

One step = one full solve of one rough-surface realisation on one GPU:
precalculation (RWG topology, octree, radiation patterns, M2L tables,
near-field impedance fill), plane-wave right-hand side and MLFMM-GMRES to
the tolerance.  Workload: 10 lambda x 10 lambda Gaussian rough patch
(RMS 0.1 lambda, correlation 0.5 lambda, seed 1), 100 x 100 cells split into
20000 triangles (N = 29,800 RWG edges, 59,600 unknowns), lambda = 633 nm,
dielectric interior n = 1.5, K = 19, s_opt = 250 (4 octree levels, 0.625
lambda leaves), normal incidence, complex128.

Paths timed (K solves each): the plain FP64 solve (the reference's
algorithm: CERFACS GMRES with IMGS on the complex128 operator, and with its
ICGS option), the mixed-precision solve (complex64 operator and inner GMRES,
FP64 iterative refinement to the same tolerance) and the near-field
preconditioned flexible GMRES (FP64 outer steps, complex64 inner GMRES on
the near-field operator).  The far-field (BRDF) intensity of each solution
is compared with the FP64 solution's; the headline value is the fastest
path whose far-field error is within 1e-4 (the FP64 path trivially is).
The batched object times 16 incidence angles solved together.

Prints one JSON line (see the repository's benchmark contract).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE on the
# default workload (100 x 100 cells, 10 lambda, ICTF, complex128), each counter
# scaled by what it reads for a plain 1 GiB device copy (FETCH x 2.0,
# WRITE x 1.0); profiles/matvec_traffic_rough10x10_ictf.txt and
# profiles/near_spmv_bf16_rough10x10_ictf.txt, profiles/near_spmv_fp8_rough10x10_ictf.txt
MEASURED_TRAFFIC = {("near_spmv", 100, 10.0, "fp64"): 1.8555e9,
                    ("aggregation", 100, 10.0, "fp64"): 1.4732e9,
                    ("disaggregation", 100, 10.0, "fp64"): 1.6627e9,
                    ("near_spmv", 100, 10.0, "c64"): 0.9289e9,
                    ("aggregation", 100, 10.0, "c64"): 0.7573e9,
                    ("disaggregation", 100, 10.0, "c64"): 0.8181e9,
                    ("near_spmv_bf16", 100, 10.0, "bf16"): 0.4716e9,
                    ("near_spmv_fp8", 100, 10.0, "fp8"): 0.2462e9}


def _sync():
    from specklesim_amd import synchronize

    synchronize(0)


def workload(args):
    from specklesim_amd import meshes

    lam = 633e-9
    pp, tt = meshes.rough_surface(args.cells, args.cells, args.size, args.size, 0.1, 0.5,
                                  args.seed, lam)
    return lam, pp, tt


class _Path:
    """One solver handle and its device vectors, kept for every step of a
    path: each step sets the handle up again from the host mesh (the whole
    precalculation runs again; its device allocations are reused, as for
    the reference's loop over surface realisations)."""

    def __init__(self, args, lam, precision, R=1):
        from specklesim_amd import DeviceBuffer, SieTriMlfmm

        self.s = SieTriMlfmm(lam, args.eps_r2, formulation=args.formulation, normal_convention=2,
                             tree_s_opt=250, truncation_number=19, precision=precision)
        self.R = R
        self.bs, self.xs = [], []
        self._DeviceBuffer = DeviceBuffer

    def vectors(self, N):
        nb = 2 * N * 16
        if not self.bs or self.bs[0].nbytes != nb:
            self.bs = [self._DeviceBuffer(nb) for _ in range(self.R)]
            self.xs = [self._DeviceBuffer(nb) for _ in range(self.R)]
        return self.bs, self.xs


def one_solve(args, path, pp, tt, orth=None, near=False):
    """Precalculation from the host mesh, then the right-hand side and GMRES
    on device buffers (b and x stay in HBM).  near: flexible GMRES
    preconditioned by inner GMRES on the near-field operator."""
    from specklesim_amd import plane_wave

    t0 = time.perf_counter()
    s = path.s
    s.precalculation_tri(pp, tt)
    t1 = time.perf_counter()
    k, e = plane_wave(0.0, 0.0, 2, "TM")
    (d_b,), (d_x,) = path.vectors(s.N)
    s.get_vector_b_tri_device(k, e, d_b.ptr)
    info = s.gmres_run_device(d_b.ptr, d_x.ptr, tolerance=args.tol,
                              max_iterations=args.maxit, orthogonalization=orth or args.orth,
                              use_initial_guess=False, near_field=near,
                              inner_iterations=args.inner)
    t2 = time.perf_counter()
    return s, d_x, dict(setup_s=t1 - t0, gmres_s=t2 - t1, total_s=t2 - t0, iters=info["iters"],
                        be=info["be"], info=info["info"], matvecs=info["matvecs"])


def spread(runs, key="total_s"):
    """min / median / max of one per-step time, in ms."""
    v = np.array([r[key] for r in runs]) * 1e3
    return {"min": round(float(v.min()), 3), "median": round(float(np.median(v)), 3),
            "max": round(float(v.max()), 3)}


def timed_path(args, lam, pp, tt, precision, orth=None, near=False):
    """W untimed solves, then exactly K solves between two device syncs."""
    path = _Path(args, lam, precision)
    for _ in range(args.warmup):
        one_solve(args, path, pp, tt, orth, near)
    _sync()
    t0 = time.perf_counter()
    runs = []
    for _ in range(args.steps):
        s, d_x, r = one_solve(args, path, pp, tt, orth, near)
        runs.append(r)
    _sync()
    elapsed = time.perf_counter() - t0
    return dict(solver=s, x=d_x, runs=runs, elapsed=elapsed,
                solves_per_s=args.steps / elapsed, ms_per_step=elapsed / args.steps * 1e3)


def batch_angles(R):
    """R polar incidence angles uniform in 0-60 degrees (the ensemble's set),
    phi = 0, TM polarisation, propagating towards -z."""
    return [np.radians(60.0 * r / max(R - 1, 1)) for r in range(R)]


def one_batch(args, path, pp, tt):
    """Precalculation, R right-hand sides on the device and one batched
    near-field preconditioned solve of all of them (FP64 outer steps,
    complex64 inner near-field steps)."""
    from specklesim_amd import plane_wave

    t0 = time.perf_counter()
    s = path.s
    s.precalculation_tri(pp, tt)
    t1 = time.perf_counter()
    bs, xs = path.vectors(s.N)
    for b, th in zip(bs, batch_angles(path.R)):
        k, e = plane_wave(th, 0.0, 2, "TM")
        s.get_vector_b_tri_device(k, e, b.ptr)
    infos = s.gmres_run_batch_device([b.ptr for b in bs], [x.ptr for x in xs],
                                     max_iterations=args.maxit, tolerance=args.tol,
                                     use_initial_guess=False, near_field=True,
                                     inner_iterations=args.inner)
    t2 = time.perf_counter()
    return s, xs, dict(setup_s=t1 - t0, gmres_s=t2 - t1, total_s=t2 - t0, infos=infos)


def timed_batch(args, lam, pp, tt, R):
    path = _Path(args, lam, "mixed", R)
    for _ in range(args.warmup):
        one_batch(args, path, pp, tt)
    _sync()
    t0 = time.perf_counter()
    runs = []
    for _ in range(args.steps):
        s, xs, r = one_batch(args, path, pp, tt)
        runs.append(r)
    _sync()
    elapsed = time.perf_counter() - t0
    return dict(solver=s, xs=xs, runs=runs, elapsed=elapsed,
                solves_per_s=R * args.steps / elapsed, ms_per_step=elapsed / args.steps * 1e3)


def ensemble_meshes(M, lam, seed0=1):
    """M seeded realisations of the speckle-ensemble statistics: patch side
    L ~ U(5, 15) lambda, RMS height ~ U(0.05, 0.15) lambda, correlation length
    ~ U(0.3, 1.0) lambda, lambda/10 cells (2 triangles each)."""
    from specklesim_amd import meshes

    rng = np.random.default_rng(12345)
    out, sizes = [], []
    for i in range(M):
        L, rms, corr = rng.uniform(5, 15), rng.uniform(0.05, 0.15), rng.uniform(0.3, 1.0)
        c = int(round(10 * L))
        out.append(meshes.rough_surface(c, c, c / 10, c / 10, rms, corr, seed0 + i, lam))
        sizes.append(c)
    return out, sizes


def timed_ensemble(args, lam):
    """The speckle ensemble: M realisations x 16 incidence angles (0-60 deg,
    TM), near-field preconditioned flexible GMRES on SPK_MIXED handles in
    batches of 16, intensity at 181 far-field points of the plane of
    incidence; worker threads with their own handles.  Then the plain FP64
    GMRES (the reference's algorithm, one solve at a time) on every
    realisation at normal incidence, for the FP64 comparison."""
    from specklesim_amd import SieTriMlfmm, ensemble_intensity, plane_wave

    ms, cells = ensemble_meshes(args.ensemble, lam)
    angles = [plane_wave(th, 0.0, 2, "TM") for th in batch_angles(16)]
    th = np.radians(np.linspace(-80, 80, 181))
    pts = 1e3 * lam * np.stack([np.sin(th), np.zeros_like(th), np.cos(th)], axis=1)
    kw = dict(formulation=args.formulation, tolerance=args.tol, max_iterations=args.maxit,
              inner_iterations=args.inner, workers=args.workers)
    ensemble_intensity(lam, args.eps_r2, ms[:2], angles, pts, **kw)  # warm-up
    _sync()
    t0 = time.perf_counter()
    I, info = ensemble_intensity(lam, args.eps_r2, ms, angles, pts, **kw)
    _sync()
    el = time.perf_counter() - t0
    its = [i["iters"] for row in info for i in row]
    # the same ensemble tuned first on realisation 0 (spk_autotune inside the
    # call: row tiles of the batched fp8 product, inner solve length)
    kw_t = dict(kw, inner_iterations="tune")
    _sync()
    t0 = time.perf_counter()
    _, info_t = ensemble_intensity(lam, args.eps_r2, ms, angles, pts, **kw_t)
    _sync()
    el_t = time.perf_counter() - t0
    # plain FP64 GMRES, one realisation at a time, normal incidence, on the
    # first (at most) 16 realisations: a per-solve rate
    s64 = SieTriMlfmm(lam, args.eps_r2, formulation=args.formulation, normal_convention=2,
                      precision="fp64")
    k, e = plane_wave(0.0, 0.0, 2, "TM")
    ms64 = ms[:16]
    t0 = time.perf_counter()
    it64 = []
    for pp, tt in ms64:
        s64.precalculation_tri(pp, tt)
        _, i64 = s64.gmres_run(s64.get_vector_b_tri(k, e), tolerance=args.tol,
                               max_iterations=args.maxit, orthogonalization=args.orth,
                               use_initial_guess=False)
        it64.append(i64["iters"])
    _sync()
    el64 = time.perf_counter() - t0
    n_edges = [3 * c * c - 2 * c for c in cells]  # interior edges of c x c cells, 2 triangles each
    return {
        "workload": f"speckle ensemble: {args.ensemble} seeded realisations (L ~ U(5,15) lambda, "
                    f"RMS ~ U(0.05,0.15) lambda, correlation ~ U(0.3,1.0) lambda, lambda/10 "
                    f"cells) x 16 incidence angles (0-60 deg, TM), {args.formulation}, n=1.5, "
                    f"tol {args.tol:g}, intensity at 181 far-field points; {args.workers} worker "
                    f"threads with their own handles (spk_ensemble_intensity)",
        "realisations": args.ensemble, "angles": 16,
        "unknowns": [2 * min(n_edges), 2 * max(n_edges)],
        "solves_per_s": len(ms) * 16 / el, "ms_per_ensemble": el * 1e3,
        "iterations": [min(its), max(its)],
        "all_converged": all(i["info"] == 0 for row in info for i in row),
        "tuned": {"solves_per_s": len(ms) * 16 / el_t, "ms_per_ensemble": el_t * 1e3,
                  "includes": "spk_autotune on realisation 0 (3 tile maps, 6 batched solves)",
                  "all_converged": all(i["info"] == 0 for row in info_t for i in row)},
        "fp64_plain_solves_per_s": len(ms64) / el64, "fp64_plain_realisations": len(ms64),
        "fp64_plain_iterations": [min(it64), max(it64)],
        "vs_fp64_plain": (len(ms) * 16 / el) / (len(ms64) / el64),
    }


def timed_adjoint(args, lam, pp, tt):
    """C5: dJ/dz of every node height of the benchmark mesh, J the sum of
    |E|^2 over a 5 x 5 detector window of +-5 degrees around the specular
    direction (normal incidence, TM, points at 1000 lambda), by one forward
    and one transposed near-field preconditioned solve (tolerance 1e-8) and
    the per-node shares (local central differences of step 1e-5 lambda) on a
    SPK_MIXED handle; the FP64 handle's gradient as
    the reference; central differences of J at --adjoint-nodes seeded
    interior nodes from FP64 solves to 1e-11, at a step of 1e-5 lambda (the
    check) and of 1e-4 lambda (the step the C5 definition names; J is not
    smooth enough at that scale for a 1e-3 check, see the record)."""
    from specklesim_amd import SieTriMlfmm, plane_wave
    from specklesim_amd.adjoint import HeightIntensity, detector_window

    k, e = plane_wave(0.0, 0.0, 2, "TM")
    pts = detector_window(0.0, 5.0, 5, 1e3 * lam)
    z0 = pp[:, 2].copy()
    nodes = np.arange(pp.shape[0])
    kw = dict(formulation=args.formulation, normal_convention=2)

    def problem(prec, tol):
        s = SieTriMlfmm(lam, args.eps_r2, precision=prec, **kw)
        return HeightIntensity(s, pp, tt, k, e, pts, tolerance=tol, dz=1e-5 * lam)

    pm = problem("mixed", 1e-8)
    pm.gradient(z0, nodes[:8])  # warm-up
    runs = []
    for _ in range(max(1, args.steps)):
        _sync()
        t0 = time.perf_counter()
        J, g, info = pm.gradient(z0, nodes)
        _sync()
        runs.append(time.perf_counter() - t0)
    t0 = time.perf_counter()
    _, g64, info64 = problem("fp64", 1e-8).gradient(z0, nodes)
    _sync()
    t64 = time.perf_counter() - t0
    # central differences at seeded interior nodes
    x, y = pp[:, 0], pp[:, 1]
    inner = np.nonzero((np.abs(x) < 0.99 * np.abs(x).max()) &
                       (np.abs(y) < 0.99 * np.abs(y).max()))[0]
    fdn = np.sort(np.random.default_rng(7).choice(inner, args.adjoint_nodes, replace=False))
    pf = problem("fp64", 1e-11)
    pf.value(z0)
    sing0 = pf.s.stats()["singular_pairs"]

    def central(h):
        fd, flips = [], 0
        t0 = time.perf_counter()
        for i in fdn:
            zp, zm = z0.copy(), z0.copy()
            zp[i] += h
            zm[i] -= h
            Jp = pf.value(zp)
            flips += int(pf.s.stats()["singular_pairs"] != sing0)
            Jm = pf.value(zm)
            flips += int(pf.s.stats()["singular_pairs"] != sing0)
            fd.append((Jp - Jm) / (2 * h))
        fd = np.array(fd)
        return {"nodes": int(len(fdn)), "step_lambda": h / lam,
                "max_err_of_max_fd": float(np.max(np.abs(g[fdn] - fd)) / np.max(np.abs(fd))),
                "branch_switches_in_stencils": flips, "forward_solves": "FP64 handle, to 1e-11",
                "seconds": time.perf_counter() - t0}

    fd5, fd4 = central(1e-5 * lam), central(1e-4 * lam)
    ms = np.array(runs) * 1e3
    return {
        "workload": f"C5: the benchmark mesh's {len(nodes)} node heights, J = sum |E|^2 over a "
                    f"5 x 5 detector window of +-5 deg around specular (normal incidence, TM), "
                    f"{args.formulation}; forward + transposed near-field preconditioned "
                    f"solves to 1e-8 and the per-node shares (spk_adjoint_height_gradient) on "
                    f"a mixed handle",
        "design_variables": len(nodes), "adjoint_step_lambda": 1e-5,
        "gradient_ms": float(np.median(ms)),
        "gradient_step_ms": {"min": float(ms.min()), "median": float(np.median(ms)),
                             "max": float(ms.max())},
        "solve_iterations": [info[0]["iters"], info[1]["iters"]],
        "J": J,
        "fp64_gradient_ms": t64 * 1e3,
        "rel_err_vs_fp64_gradient": float(np.linalg.norm(g - g64) / np.linalg.norm(g64)),
        "fd_check": fd5,
        "fd_at_1e-4_lambda": fd4,
    }


def brdf_curve(res, lam):
    """BRDF of a solution in the plane of incidence, 181 angles in
    (-80, 80) degrees on a sphere of 1000 wavelengths."""
    from specklesim_amd import farfield

    _, v = farfield.brdf(res["solver"], res["x"].ptr, 0.0, (-np.radians(80), np.radians(80)), 181,
                         0.0, 1e3 * lam, device=True)
    return v


def cpu_baseline(args, lam, pp, tt, iters):
    """The restatement (oracle/) on the host cores, timed on a bounded sample
    of one solve and scaled to the whole solve: the near fill of 2 % of the
    rows (scaled by nnz), the tree, tables and patterns (whole), three
    matvecs, and a 30-step CERFACS IMGS GMRES (matvecs plus Gram-Schmidt),
    whose orthogonalisation time, linear in the step index, is scaled to the
    FP64 path's iteration count."""
    from oracle.oracle import Oracle

    nthreads = min(16, os.cpu_count() or 1)
    o = Oracle(pp, tt, lam, args.eps_r2, formulation=args.formulation, normal_convention=2)
    t0 = time.perf_counter()
    o.tree()
    ia, ja, _ = o.near_csr(fill=False)
    t_tree = time.perf_counter() - t0
    rows = max(1, o.N // 50)
    t0 = time.perf_counter()
    o.near_fill_rows(0, rows, nthreads)
    t_fill_sample = time.perf_counter() - t0
    t_fill = t_fill_sample * (ia[-1] / max(1, ia[rows]))
    t0 = time.perf_counter()
    o.mlfmm_setup(nthreads)
    t_mlfmm_setup = time.perf_counter() - t0
    rng = np.random.default_rng(7)
    x = rng.standard_normal(2 * o.N) + 1j * rng.standard_normal(2 * o.N)
    o.matvec(x, nthreads)
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        o.matvec(x, nthreads)
    t_mv = (time.perf_counter() - t0) / reps
    # 30 GMRES steps: matvecs + IMGS orthogonalisation against j vectors at
    # step j, so the Gram-Schmidt time of `iters` steps is (iters/30)^2 of it
    ks = 30
    b = o.rhs_plane(np.array([0.0, 0.0, -1.0]), np.array([1.0, 0.0, 0.0]))
    t0 = time.perf_counter()
    o.gmres(b, tol=1e-30, maxit=ks, orth=1, nthreads=nthreads)
    t_g = time.perf_counter() - t0
    t_orth = max(0.0, t_g - (ks + 1) * t_mv)
    t_solve = (t_tree + t_fill + t_mlfmm_setup + (iters + 2) * t_mv
               + t_orth * (iters / ks) ** 2)
    return {"value": 1.0 / t_solve, "unit": "solves/s", "cores": nthreads, "kind": "port",
            "sample": f"oracle/ restatement, {nthreads} OpenMP threads: near fill of rows "
                      f"[0,{rows}) scaled by nnz ({t_fill:.1f} s est.), tree+tables+patterns "
                      f"{t_tree + t_mlfmm_setup:.1f} s, {reps} matvecs ({t_mv * 1e3:.0f} ms each) x "
                      f"{iters + 2} GMRES matvecs of the FP64 GPU solve, and a {ks}-step IMGS "
                      f"GMRES ({t_g:.1f} s) whose Gram-Schmidt part ({t_orth:.2f} s) is scaled "
                      f"by ({iters}/{ks})^2; whole solve {t_solve:.1f} s est.",
            "matvec_ms": t_mv * 1e3, "solve_s_est": t_solve}


def timed_quad(args, lam):
    """C3: the MLFMM operator on a 20 x 20 lambda rough patch of 200 x 200
    8-node quadrilaterals (10 edge functions per element, M = 240,800,
    481,600 unknowns; RMS 0.1 lambda, correlation 0.5 lambda, seed 1) on a
    mixed-precision handle, ICTF, K = 19, leaf boxes of 0.625 lambda
    (tree_s_opt 400; the default 250 gives 0.3125 lambda leaves).  Setup
    (near fill from the quadrilateral pair integrals, patterns, tables),
    FP64 and complex64 products timed with HIP events, and 40 near-field
    preconditioned outer steps.  The reference solves this system by dense
    LU only; its condition number (~1e10 at 0.1 lambda elements, measured on
    small grids) stalls GMRES, preconditioned or not."""
    import gc

    from specklesim_amd import DeviceBuffer, SieTriMlfmm, meshes, plane_wave

    n = args.quad_cells
    L = n / 10.0
    Lx, Ly, nx, ny, z = meshes.rough_height_field(n, n, L, L, 0.1, 0.5, 1, lam)
    gc.collect()
    _sync()
    t0 = time.perf_counter()
    g = SieTriMlfmm(lam, args.eps_r2, formulation=args.formulation, normal_convention=2,
                    tree_s_opt=400, precision="mixed").precalculation_quad(Lx, Ly, nx, ny, z)
    _sync()
    setup_s = time.perf_counter() - t0
    st = g.stats()
    reps = max(3, args.steps)
    p64 = g.profile_matvec(reps=reps)
    p32 = g.profile_matvec(reps=reps, c64=True)
    m64, m32 = p64["matvec"][0], p32["matvec"][0]
    # algorithmic bytes of one product: near blocks, both pattern sets, x / y
    nbytes = st["bytes_near"] + st["bytes_patterns"] + 3 * 2 * g.N * 16.0
    k, e = plane_wave(np.deg2rad(10.0), 0.0, 2, "TM")
    nb = 2 * g.N * 16
    db, dx = DeviceBuffer(nb), DeviceBuffer(nb)
    g.get_vector_b_tri_device(k, e, db.ptr)
    _sync()
    t0 = time.perf_counter()
    info = g.gmres_run_device(db.ptr, dx.ptr, tolerance=args.tol, max_iterations=40,
                              orthogonalization="ICGS", use_initial_guess=False, near_field=True)
    _sync()
    solve_s = time.perf_counter() - t0
    rec = {
        "workload": f"C3: {n} x {n} 8-node quadrilaterals on a {L:g} x {L:g} lambda rough patch "
                    f"(RMS 0.1, correlation 0.5 lambda), {args.formulation}, K = 19, leaf boxes "
                    f"of 0.625 lambda, mixed-precision handle",
        "edge_functions": g.N, "unknowns": 2 * g.N, "near_blocks": st["near_blocks"],
        "levels": [st["l_min"], st["l_max"]],
        "setup_s": setup_s, "near_fill_ms": st["near_fill_ms"], "patterns_ms": st["patterns_ms"],
        "matvec_ms_fp64": m64, "matvec_ms_c64": m32,
        "matvec_stage_ms_fp64": {kk: round(v[0], 4) for kk, v in p64.items() if kk != "matvec"},
        "fp64_product_bytes": nbytes, "fp64_product_tb_per_s": nbytes / (m64 * 1e-3) / 1e12,
        "operator_bytes_fp64_and_c64": 1.5 * (st["bytes_near"] + st["bytes_patterns"]),
        "preconditioned_outer_steps": {"steps": info["iters"], "backward_error": info["be"],
                                       "converged": info["info"] == 0, "seconds": solve_s},
    }
    del g, db, dx
    gc.collect()
    return rec


def timed_quad_dense(args, lam):
    """The reference's own quadrilateral solve (lib_sie_quad_solver_normal_I,
    src/lib_sie_quad_solver_normal.f90:80-229): the dense 2M x 2M system
    filled on the GPU, the plane-wave right-hand side and rocSOLVER's FP64
    LU (ZGESV in the reference, src/lib_sie_math.f90:649-676), on an n x n
    rough patch of 0.1 lambda quadrilaterals (RMS 0.1, correlation 0.5
    lambda, seed 1).  The factorisation is priced at 8/3 (2M)^3 real flops
    against the FP64 matrix-core peak."""
    import gc

    from specklesim_amd import DeviceBuffer, QuadFill, meshes, plane_wave

    n = args.quad_dense_cells
    L = n / 10.0
    Lx, Ly, nx, ny, z = meshes.rough_height_field(n, n, L, L, 0.1, 0.5, 1, lam)
    g = QuadFill(lam, args.eps_r2, formulation=args.formulation).setup(Lx, Ly, nx, ny, z)
    N = 2 * g.M
    MM, b = DeviceBuffer(N * N * 16), DeviceBuffer(N * 16)
    k, e = plane_wave(np.deg2rad(10.0), 0.0, 2, "TM")
    runs = []
    for _ in range(2):
        _sync()
        t0 = time.perf_counter()
        g.fill_device(MM.ptr)
        _sync()
        t1 = time.perf_counter()
        g.get_vector_b_quad_device(k, e, b.ptr)
        g.solve_lu_device(MM.ptr, b.ptr)
        _sync()
        t2 = time.perf_counter()
        runs.append({"fill_s": t1 - t0, "lu_s": t2 - t1, "total_s": t2 - t0})
    best = min(runs, key=lambda r: r["total_s"])
    flops = 8.0 / 3.0 * float(N) ** 3
    x = b.download(np.complex128)
    rec = {
        "workload": f"{n} x {n} 8-node quadrilaterals on a {L:g} x {L:g} lambda rough patch "
                    f"(RMS 0.1, correlation 0.5 lambda), {args.formulation}: dense fill, "
                    f"plane-wave right-hand side and FP64 LU solve, as the reference solves it",
        "edge_functions": g.M, "unknowns": N, "matrix_gb": N * N * 16 / 1e9,
        "fill_s": best["fill_s"], "lu_s": best["lu_s"], "solve_s": best["total_s"],
        "solves_per_s": 1.0 / best["total_s"], "runs": runs,
        "lu_roofline": {"bound": "mfma", "achieved": flops / best["lu_s"] / 1e12, "peak": 78.6,
                        "unit": "TFLOP/s", "frac": flops / best["lu_s"] / 1e12 / 78.6,
                        "flops": flops},
        "solution_finite": bool(np.isfinite(x).all()),
    }
    del MM, b, g
    gc.collect()
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=100)
    ap.add_argument("--size", type=float, default=10.0)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--eps-r2", type=complex, default=2.25 + 0j)
    ap.add_argument("--formulation", default="ICTF")
    ap.add_argument("--precision", default="both", choices=["both", "fp64", "mixed"])
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--maxit", type=int, default=650)
    ap.add_argument("--orth", default="IMGS")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch", type=int, default=16,
                    help="incidence angles of the batched path (0: skip it)")
    ap.add_argument("--inner", type=int, default=30,
                    help="inner near-field GMRES steps per flexible GMRES step")
    ap.add_argument("--ensemble", type=int, default=64,
                    help="realisations of the speckle-ensemble object (0: skip it)")
    ap.add_argument("--adjoint-nodes", type=int, default=32,
                    help="nodes of the adjoint object's finite-difference check (0: skip it)")
    ap.add_argument("--quad-cells", type=int, default=200,
                    help="quadrilaterals per side of the C3 object (0: skip it)")
    ap.add_argument("--quad-dense-cells", type=int, default=40,
                    help="quadrilaterals per side of the dense LU object (0: skip it)")
    ap.add_argument("--workers", type=int, default=3,
                    help="worker threads (handles) of the ensemble")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    lam, pp, tt = workload(args)
    paths = {}
    for prec in (("fp64", "mixed") if args.precision == "both" else (args.precision,)):
        paths[prec] = timed_path(args, lam, pp, tt, prec)
    if args.precision == "both" and args.orth != "ICGS":
        # the same FP64 solve with the reference's ICGS option (icntl(5) = 3)
        paths["fp64_icgs"] = timed_path(args, lam, pp, tt, "fp64", "ICGS")
    if args.precision == "both":
        # flexible GMRES on the FP64 operator, complex64 near-field inner solves
        paths["near_fgmres"] = timed_path(args, lam, pp, tt, "mixed", "ICGS", near=True)
        # the same preconditioned algorithm on an FP64 handle (complex128 inner
        # solves on the FP64 near blocks): separates the gain of the algorithm
        # from the gain of the low-precision copies
        paths["near_fgmres_fp64"] = timed_path(args, lam, pp, tt, "fp64", "ICGS", near=True)
        # the same as near_fgmres with the inner solve's kernels launched one
        # by one instead of replaying its captured graph
        os.environ["SPK_INNER_GRAPH"] = "0"
        try:
            paths["near_fgmres_eager"] = timed_path(args, lam, pp, tt, "mixed", "ICGS", near=True)
        finally:
            del os.environ["SPK_INNER_GRAPH"]
    batch = batch_eager = None
    if args.batch and args.precision == "both":
        batch = timed_batch(args, lam, pp, tt, args.batch)
        os.environ["SPK_INNER_GRAPH"] = "0"
        try:
            batch_eager = timed_batch(args, lam, pp, tt, args.batch)
        finally:
            del os.environ["SPK_INNER_GRAPH"]
    ens = timed_ensemble(args, lam) if args.ensemble and args.precision == "both" else None
    adj = (timed_adjoint(args, lam, pp, tt) if args.adjoint_nodes and args.precision == "both"
           else None)
    quad = timed_quad(args, lam) if args.quad_cells and args.precision == "both" else None
    qdense = (timed_quad_dense(args, lam) if args.quad_dense_cells and args.precision == "both"
              else None)
    ff_err = None
    if "fp64" in paths and "mixed" in paths:
        b64, b32 = brdf_curve(paths["fp64"], lam), brdf_curve(paths["mixed"], lam)
        ff_err = float(np.linalg.norm(b32 - b64) / np.linalg.norm(b64))
    # far-field (BRDF) error of every path against the FP64 path's solution
    ff = {k: float(np.linalg.norm(brdf_curve(v, lam) - b64) / np.linalg.norm(b64))
          for k, v in paths.items() if k != "fp64"} if "fp64" in paths else {}
    # the headline is the fastest product path within the far-field bound;
    # near_fgmres_eager (the same solver with its inner solve launched
    # kernel by kernel) is a diagnostic variant, not a candidate
    ok = {k: v for k, v in paths.items()
          if k == "fp64" or (not k.endswith("_eager") and ff.get(k, 1.0) <= 1e-4
                             and v["runs"][-1]["info"] == 0)}
    head = max(ok, key=lambda k: ok[k]["solves_per_s"]) if ok else next(iter(paths))
    H = paths[head]
    runs, elapsed, ms_per_step = H["runs"], H["elapsed"], H["ms_per_step"]
    s = paths["fp64"]["solver"] if "fp64" in paths else H["solver"]

    prof = s.profile_matvec(reps=20)
    st = s.stats()
    # the roofline kernel: the longest HBM-bound stage of the headline path's
    # matvec (the complex64 operator when the mixed path leads)
    pc = paths["mixed"]["solver"].profile_matvec(reps=20, c64=True) if "mixed" in paths else None
    if head == "near_fgmres":
        # the near-field preconditioner's fp8 SpMV (one right-hand side): 30
        # launches per outer step, the longest-running kernel of this path
        stage, rprec = "near_spmv_fp8", "fp8"
        ms, by = H["solver"].profile_near_bf(reps=20, fp8=True)
    else:
        rprof, rprec = (pc, "c64") if head == "mixed" else (prof, "fp64")
        stage, (ms, by) = max(((k, v) for k, v in rprof.items() if k != "matvec"),
                              key=lambda kv: kv[1][0])
    achieved = by / (ms * 1e-3) / 1e9
    traffic = MEASURED_TRAFFIC.get((stage, args.cells, args.size, rprec))
    from specklesim_amd._native import read_bandwidth

    ceiling = read_bandwidth(0, 2048, 5)
    rec = {
        "metric": "solves/s (MLFMM-GMRES to tolerance, incl. precalculation); matvec ms",
        "value": args.steps / elapsed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128" if head == "fp64" else "complex64+complex128",
        "data": "synthetic (seeded Gaussian rough surface)",
        "config": {"workload": f"rough patch {args.size:g}x{args.size:g} lambda, "
                               f"{args.cells}x{args.cells} cells, {2 * args.cells ** 2} triangles, "
                               f"N={st['n_edges']} RWG, {args.formulation}, n=1.5, K=19, "
                               f"s_opt=250, tol {args.tol:g}, {args.orth}, headline path {head}",
                   "unknowns": 2 * st["n_edges"], "near_blocks": st["near_blocks"],
                   "levels": [st["l_min"], st["l_max"]], "parallelism": "single GPU"},
        "matvec_ms": prof["matvec"][0],
        "stages_ms": {k: round(v[0], 4) for k, v in prof.items()},
        "stages_GBps": {k: round(v[1] / (v[0] * 1e-3) / 1e9, 1) for k, v in prof.items()
                        if v[0] > 0},
        "setup_ms": st["setup_ms"], "near_fill_ms": st["near_fill_ms"],
        "setup_split_ms": {k: round(st[k], 3) for k in ("tree_ms", "near_fill_ms", "patterns_ms",
                                                         "tables_ms")},
        "gmres_iterations": runs[-1]["iters"], "gmres_backward_error": runs[-1]["be"],
        "gmres_info": runs[-1]["info"], "gmres_s": runs[-1]["gmres_s"],
        "paths": {k: {"solves_per_s": v["solves_per_s"], "ms_per_step": v["ms_per_step"],
                      "iterations": v["runs"][-1]["iters"], "matvecs": v["runs"][-1]["matvecs"],
                      "backward_error": v["runs"][-1]["be"], "info": v["runs"][-1]["info"],
                      "setup_ms": v["runs"][-1]["setup_s"] * 1e3,
                      "gmres_ms": v["runs"][-1]["gmres_s"] * 1e3,
                      "step_ms": spread(v["runs"]), "setup_step_ms": spread(v["runs"], "setup_s"),
                      "gmres_step_ms": spread(v["runs"], "gmres_s")}
                  for k, v in paths.items()},
        "farfield_rel_err_mixed_vs_fp64": ff_err,
        "farfield_rel_err_vs_fp64": ff,
        "headline_path": head,
        "mixed_vs_fp64": (paths["mixed"]["solves_per_s"] / paths["fp64"]["solves_per_s"]
                          if "mixed" in paths and "fp64" in paths else None),
        # the same preconditioned flexible GMRES with low-precision inner solves
        # (complex64 vectors, fp8 near blocks) against complex128 inner solves
        "near_fgmres_mixed_vs_fp64": (paths["near_fgmres"]["solves_per_s"]
                                      / paths["near_fgmres_fp64"]["solves_per_s"]
                                      if "near_fgmres_fp64" in paths else None),
        "roofline": {"kernel": stage,
                     "operator": {"c64": "complex64", "fp64": "complex128",
                                  "bf16": "bf16 near blocks, complex64 vectors",
                                  "fp8": "fp8 (E4M3FN) near blocks with per-row component "
                                         "scales, complex64 vectors"}[rprec],
                     "bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "algorithmic_bytes": by, "ms": ms,
                     "read_ceiling": ceiling, "frac_of_read_ceiling": achieved / ceiling},
    }
    if pc is not None:
        rec["matvec_c64_ms"] = pc["matvec"][0]
        rec["stages_c64_ms"] = {k: round(v[0], 4) for k, v in pc.items()}
    if batch is not None:
        from specklesim_amd import farfield

        # far field of the batch's normal-incidence solution against the FP64
        # path's (the same right-hand side), and of its 60-degree solution
        # against an FP64 solve of that right-hand side (not timed)
        bres = dict(solver=batch["solver"], x=batch["xs"][0])
        ff_b0 = float(np.linalg.norm(brdf_curve(bres, lam) - b64) / np.linalg.norm(b64))
        from specklesim_amd import DeviceBuffer, SieTriMlfmm, plane_wave

        th = batch_angles(args.batch)[-1]
        s64 = paths["fp64"]["solver"]
        k, e = plane_wave(th, 0.0, 2, "TM")
        db = DeviceBuffer(2 * s64.N * 16)
        dx = DeviceBuffer(2 * s64.N * 16)
        s64.get_vector_b_tri_device(k, e, db.ptr)
        s64.gmres_run_device(db.ptr, dx.ptr, tolerance=args.tol, max_iterations=args.maxit,
                             use_initial_guess=False)
        pts = (-np.radians(80), np.radians(80))
        _, v64 = farfield.brdf(s64, dx.ptr, th, pts, 181, 0.0, 1e3 * lam, device=True)
        _, vb = farfield.brdf(batch["solver"], batch["xs"][-1].ptr, th, pts, 181, 0.0, 1e3 * lam,
                              device=True)
        ff_b1 = float(np.linalg.norm(vb - v64) / np.linalg.norm(v64))
        infos = batch["runs"][-1]["infos"]
        rec["batched"] = {
            "workload": f"the same mesh, {args.batch} incidence angles uniform in 0-60 deg "
                        f"(TM, phi = 0), solved together: flexible GMRES with batched FP64 "
                        f"products, preconditioned by {args.inner} batched complex64 "
                        f"near-field GMRES steps",
            "solves_per_s": batch["solves_per_s"], "ms_per_batch": batch["ms_per_step"],
            "right_hand_sides": args.batch,
            "setup_ms": batch["runs"][-1]["setup_s"] * 1e3,
            "gmres_ms": batch["runs"][-1]["gmres_s"] * 1e3,
            "step_ms": spread(batch["runs"]), "setup_step_ms": spread(batch["runs"], "setup_s"),
            "gmres_step_ms": spread(batch["runs"], "gmres_s"),
            "iterations": [min(i["iters"] for i in infos), max(i["iters"] for i in infos)],
            "backward_error_max": max(i["be"] for i in infos),
            "all_converged": all(i["info"] == 0 for i in infos),
            "farfield_rel_err_vs_fp64": {"0deg": ff_b0, f"{np.degrees(th):.0f}deg": ff_b1},
            "vs_fp64_single": batch["solves_per_s"] / paths["fp64"]["solves_per_s"],
            "vs_mixed_single": batch["solves_per_s"] / paths["mixed"]["solves_per_s"],
            "vs_headline_single": batch["solves_per_s"] / H["solves_per_s"],
            "inner_graph": "captured inner near-field solve replayed per outer step",
            "eager_ms_per_batch": batch_eager["ms_per_step"],
            "eager_step_ms": spread(batch_eager["runs"]),
        }
    if ens is not None:
        rec["ensemble"] = ens
    if adj is not None:
        rec["adjoint"] = adj
    if quad is not None:
        rec["quad_mlfmm"] = quad
    if qdense is not None:
        rec["quad_dense_lu"] = qdense
    if not args.no_cpu_baseline:
        it64 = (paths["fp64"] if "fp64" in paths else H)["runs"][-1]["iters"]
        cb = cpu_baseline(args, lam, pp, tt, it64)
        rec["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        rec["cpu_matvec_ms"] = cb["matvec_ms"]
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
