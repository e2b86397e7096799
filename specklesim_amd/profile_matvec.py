"""Matvec profiling driver: set up one rough-surface mesh and repeat the
MLFMM + near-field matvec, printing per-stage device times (HIP events on
the handle's stream) and algorithmic HBM bytes as one JSON line.  Meant to
run under ``rocprofv3`` (kernel trace or counters) without the GMRES solve
around it.

    python -m specklesim_amd.profile_matvec [--cells 100] [--size 10] [--reps 20]
    python -m specklesim_amd.profile_matvec --copy-mib 1024
    python -m specklesim_amd.profile_matvec --batch 16
    python -m specklesim_amd.profile_matvec --near-bf
    python -m specklesim_amd.profile_matvec --near-f8
    python -m specklesim_amd.profile_matvec --solve near

``--copy-mib`` instead runs five plain device-to-device copies of that size
(torch's copy kernel), the calibration for FETCH_SIZE / WRITE_SIZE counters.
``--near-bf`` instead repeats the near-field preconditioner's bf16 SpMV,
``--near-f8`` its fp8 (E4M3FN) SpMV.
``--quad N`` instead sets up the MLFMM on an N x N quadrilateral patch and
times its products (``--precision mixed`` for the complex64 mirror).
``--quad-dense N`` instead fills the dense system of an N x N quadrilateral
patch and solves it by LU (the reference's quadrilateral solve), twice.
``--solve near|mixed|fp64`` instead runs one solve of the normal-incidence
TM plane wave to 1e-6 (near: the near-field preconditioned flexible GMRES;
mixed: the refinement solve; fp64: the reference's GMRES with IMGS).
``--batch R`` instead runs one batched near-field preconditioned solve of R plane
waves (polar angles uniform in 0-60 degrees, TM, tolerance 1e-6) after the
precalculation, and prints each right-hand side's iterations and backward
error.
"""
from __future__ import annotations

import argparse
import json


def copy_calibration(mib: int) -> dict:
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    return {"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 5}


def batched_solve(s, R: int) -> dict:
    import numpy as np

    from . import DeviceBuffer, plane_wave, synchronize

    nb = 2 * s.N * 16
    bs = [DeviceBuffer(nb) for _ in range(R)]
    xs = [DeviceBuffer(nb) for _ in range(R)]
    for r, b in enumerate(bs):
        k, e = plane_wave(np.radians(60.0 * r / max(R - 1, 1)), 0.0, 2, "TM")
        s.get_vector_b_tri_device(k, e, b.ptr)
    infos = s.gmres_run_batch_device([b.ptr for b in bs], [x.ptr for x in xs],
                                     max_iterations=650, tolerance=1e-6,
                                     use_initial_guess=False, near_field=True)
    synchronize(0)
    return {"batch": R, "iters": [i["iters"] for i in infos], "be": [i["be"] for i in infos]}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--cells", type=int, default=100)
    ap.add_argument("--size", type=float, default=10.0)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--eps-r2", type=complex, default=2.25 + 0j)
    ap.add_argument("--formulation", default="ICTF")
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--copy-mib", type=int, default=0)
    ap.add_argument("--batch", type=int, default=0)
    ap.add_argument("--near-bf", action="store_true")
    ap.add_argument("--near-f8", action="store_true")
    ap.add_argument("--solve", default="", choices=["", "near", "mixed", "fp64"])
    ap.add_argument("--adjoint", action="store_true",
                    help="one adjoint gradient over every node height (detector window of "
                         "+-5 deg around specular, solves to 1e-8)")
    ap.add_argument("--quad", type=int, default=0,
                    help="the MLFMM products on an N x N quadrilateral rough patch of N/10 "
                         "lambda (RMS 0.1, correlation 0.5 lambda, tree_s_opt 400) instead")
    ap.add_argument("--quad-dense", type=int, default=0,
                    help="dense fill and LU solve of an N x N quadrilateral rough patch of N/10 "
                         "lambda instead")
    args = ap.parse_args(argv)
    if args.copy_mib:
        print(json.dumps(copy_calibration(args.copy_mib)), flush=True)
        return

    from . import SieTriMlfmm, meshes

    lam = 633e-9
    if args.quad_dense:
        import time

        from . import DeviceBuffer, QuadFill, plane_wave
        from ._native import synchronize
        n = args.quad_dense
        Lx, Ly, nx, ny, z = meshes.rough_height_field(n, n, n / 10.0, n / 10.0, 0.1, 0.5,
                                                      args.seed, lam)
        g = QuadFill(lam, args.eps_r2, formulation=args.formulation).setup(Lx, Ly, nx, ny, z)
        N = 2 * g.M
        MM, b = DeviceBuffer(N * N * 16), DeviceBuffer(N * 16)
        k, e = plane_wave(0.0, 0.0, 2, "TM")
        out = []
        for _ in range(2):
            t0 = time.perf_counter()
            g.fill_device(MM.ptr)
            t1 = time.perf_counter()
            g.get_vector_b_quad_device(k, e, b.ptr)
            g.solve_lu_device(MM.ptr, b.ptr)
            synchronize()
            out.append({"fill_s": t1 - t0, "lu_s": time.perf_counter() - t1})
        print(json.dumps({"workload": f"quadrilateral rough patch {n / 10:g} lambda, {n}x{n} "
                                      f"elements, dense {N} x {N} LU, {args.formulation}",
                          "runs": out}), flush=True)
        return
    if args.quad:
        n = args.quad
        Lx, Ly, nx, ny, z = meshes.rough_height_field(n, n, n / 10.0, n / 10.0, 0.1, 0.5,
                                                      args.seed, lam)
        s = SieTriMlfmm(lam, args.eps_r2, formulation=args.formulation, normal_convention=2,
                        tree_s_opt=400, truncation_number=19, precision=args.precision)
        s.precalculation_quad(Lx, Ly, nx, ny, z)
        st = s.stats()
        prof = s.profile_matvec(reps=args.reps, c64=args.precision == "mixed")
        print(json.dumps({"workload": f"quadrilateral rough patch {n / 10:g} lambda, {n}x{n} "
                                      f"elements, M={st['n_edges']}, {args.formulation}",
                          "operator": "complex64" if args.precision == "mixed" else "complex128",
                          "setup_ms": st["setup_ms"], "near_fill_ms": st["near_fill_ms"],
                          "bytes_near": st["bytes_near"], "bytes_patterns": st["bytes_patterns"],
                          "stages_ms": {k: v[0] for k, v in prof.items()}}), flush=True)
        return
    pp, tt = meshes.rough_surface(args.cells, args.cells, args.size, args.size, 0.1, 0.5,
                                  args.seed, lam)
    s = SieTriMlfmm(lam, args.eps_r2, formulation=args.formulation, normal_convention=2,
                    tree_s_opt=250, truncation_number=19,
                    precision=("fp64" if args.solve == "fp64" else "mixed")
                    if (args.batch or args.near_bf or args.near_f8 or args.solve or args.adjoint)
                    else args.precision)
    if args.adjoint:
        import time

        import numpy as np

        from . import plane_wave
        from .adjoint import HeightIntensity, detector_window

        k, e = plane_wave(0.0, 0.0, 2, "TM")
        prob = HeightIntensity(s, pp, tt, k, e, detector_window(0.0, 5.0, 5, 1e3 * lam),
                               tolerance=1e-8, dz=1e-6 * lam)
        t0 = time.perf_counter()
        J, _, info = prob.gradient(pp[:, 2], np.arange(pp.shape[0]))
        print(json.dumps({"adjoint": pp.shape[0], "J": J, "info": info,
                          "seconds": time.perf_counter() - t0}), flush=True)
        return
    s.precalculation_tri(pp, tt)
    if args.batch:
        print(json.dumps(batched_solve(s, args.batch)), flush=True)
        return
    if args.solve:
        from . import plane_wave

        k, e = plane_wave(0.0, 0.0, 2, "TM")
        b = s.get_vector_b_tri(k, e)
        x, inf = s.gmres_run(b, tolerance=1e-6, max_iterations=650, use_initial_guess=False,
                             near_field=args.solve == "near")
        print(json.dumps({"solve": args.solve, "iters": inf["iters"], "matvecs": inf["matvecs"],
                          "be": inf["be"], "seconds": inf["seconds"]}), flush=True)
        return
    if args.near_bf or args.near_f8:
        ms, by = s.profile_near_bf(reps=args.reps, fp8=args.near_f8)
        print(json.dumps({"kernel": "k_near_spmv_f8" if args.near_f8 else "k_near_spmv_bf",
                          "reps": args.reps, "ms": ms, "bytes": by,
                          "GBps": by / (ms * 1e-3) / 1e9}), flush=True)
        return
    c64 = args.precision == "mixed"
    prof = s.profile_matvec(reps=args.reps, c64=c64)
    st = s.stats()
    out = {"workload": f"rough patch {args.size:g}x{args.size:g} lambda, {args.cells}x{args.cells} "
                       f"cells, N={st['n_edges']}, {args.formulation}",
           "operator": "complex64" if c64 else "complex128",
           "reps": args.reps, "near_blocks": st["near_blocks"],
           "stages": {k: {"ms": v[0], "bytes": v[1],
                          "GBps": (v[1] / (v[0] * 1e-3) / 1e9) if v[0] > 0 else None}
                      for k, v in prof.items()}}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
