"""Result files of a GMRES solve in the reference's text formats, so that a
run of this library can be compared line by line with a reference run.

* ``Result_SEH_*``: ``lib_sie_solver_gmres_run``
  (src/lib_sie_solver_gmres_interface.f90:163-170) writes ``n_el`` and
  ``m_pairs`` list-directed, then the 2N coefficients [xe; xh], one complex
  per line as ``(es19.12, tr5)`` for the real and the imaginary part.
* ``iteration_error_*``: ``zgmres_backward_error`` (:82-92) writes
  ``(i10, tr5, es19.12, tr5)`` = iteration, Arnoldi backward error.  The
  solver calls it only where it tests convergence, i.e. on the iterations
  whose Arnoldi backward error is at or below the tolerance, and on the last
  allowed iteration (src/lib/libmath_external/src/algorithm_842/
  zPackgmres.f90:1337-1352).  The file name gets ``_LeftP.txt`` or
  ``_NoP.txt`` by preconditioning (lib_sie_solver_gmres_interface.f90:142-146).
* file names follow ``name_outputfile_multiple`` / ``name_outputfile_single``
  (main.f90:625-705).
"""
from __future__ import annotations

import math

import numpy as np


def fortran_es(v: float, w: int = 19, d: int = 12) -> str:
    """``ESw.d`` edit of one real: one non-zero digit before the point, a
    two-digit exponent with its letter (``E+05``) or, beyond 99, a
    three-digit exponent without it (``-100``), right-justified in w."""
    v = float(v)
    if not math.isfinite(v):
        s = "NaN" if math.isnan(v) else ("-Infinity" if v < 0 else "Infinity")
        return s.rjust(w)
    m = f"{abs(v):.{d}E}"
    mant, ex = m.split("E")
    e = int(ex)
    exp = f"E{'-' if e < 0 else '+'}{abs(e):02d}" if abs(e) <= 99 else f"{'-' if e < 0 else '+'}{abs(e):03d}"
    s = ("-" if math.copysign(1.0, v) < 0 and v != 0 else "") + mant + exp
    if len(s) > w:
        return "*" * w
    return s.rjust(w)


def _parse_fortran_real(tok: str) -> float:
    t = tok.strip()
    try:
        return float(t)
    except ValueError:
        # exponent without its letter: "1.0-100" / "1.0+100"
        for i in range(len(t) - 1, 0, -1):
            if t[i] in "+-" and t[i - 1] not in "eE":
                return float(t[:i] + "E" + t[i:])
        raise


def write_result_seh(path, n_el: int, vector_x) -> None:
    """The solution file (lib_sie_solver_gmres_interface.f90:163-168)."""
    x = np.asarray(vector_x, dtype=np.complex128).ravel()
    if x.size % 2:
        raise ValueError("vector_x holds [xe; xh]: its length must be even")
    with open(path, "w") as f:
        f.write(f" n_el,{int(n_el):12d}\n")
        f.write(f" m_pairs{x.size // 2:12d}\n")
        for c in x:
            f.write(fortran_es(c.real) + " " * 5 + fortran_es(c.imag) + "\n")


def read_result_seh(path):
    """(n_el, m_pairs, vector_x) from a solution file."""
    with open(path) as f:
        h1 = f.readline().replace(",", " ").split()
        h2 = f.readline().replace(",", " ").split()
        n_el, m_pairs = int(h1[-1]), int(h2[-1])
        vals = []
        for line in f:
            p = line.split()
            if not p:
                continue
            if len(p) != 2:
                raise ValueError(f"expected two reals per line, got {line!r}")
            vals.append(complex(_parse_fortran_real(p[0]), _parse_fortran_real(p[1])))
    x = np.array(vals, dtype=np.complex128)
    if x.size != 2 * m_pairs:
        raise ValueError(f"{path}: {x.size} coefficients for m_pairs = {m_pairs}")
    return n_el, m_pairs, x


def convergence_checks(history, tolerance: float, max_iterations: int):
    """The (iteration, Arnoldi backward error) pairs the reference's callback
    receives: iterations (1-based) whose backward error is <= tolerance, and
    the last one when the solve stops at the iteration limit."""
    h = np.asarray(history, dtype=np.float64)
    out = []
    for k, b in enumerate(h):
        it = k + 1
        if b <= tolerance or it >= max_iterations:
            out.append((it, float(b)))
    return out


def write_iteration_error(path, history, tolerance: float, max_iterations: int) -> None:
    """The convergence file (lib_sie_solver_gmres_interface.f90:91)."""
    with open(path, "w") as f:
        for it, b in convergence_checks(history, tolerance, max_iterations):
            f.write(f"{it:10d}" + " " * 5 + fortran_es(b) + "\n")


def read_iteration_error(path):
    its, bes = [], []
    with open(path) as f:
        for line in f:
            p = line.split()
            if p:
                its.append(int(p[0]))
                bes.append(_parse_fortran_real(p[1]))
    return np.array(its, dtype=np.int64), np.array(bes)


def output_names(formulation: str, index: int | None = None, element: str = "tri",
                 evaluation: str = "BRDF_n", preconditioned: bool = False,
                 multi_mesh: bool = False) -> dict:
    """File names of one GMRES run (main.f90:625-705): ``index`` set names
    the numbered run of an ensemble (``I3.3``), else a single run;
    multi_mesh: the ensemble reads numbered pp/tt meshes (calc_p(1) = 4), whose
    result name carries the evaluation type.  formulation is one of PMCHWT,
    MCTF, ICTF, MCTF2, evaluation one of xz, xy, yz, rcs_n, rcs_p, BRDF_n,
    BRDF_p, cyl_field (main.f90:588-616)."""
    form = formulation.strip()
    suffix = "_LeftP.txt" if preconditioned else "_NoP.txt"
    if index is None:
        # a single run builds the names in character(len = 10) variables
        # (main.f90:685), which cut "cyl_field.txt" to "cyl_field."
        names = {"result": f"Result_SEH_{element}_" + f"{form}.txt"[:10],
                 "efield": f"Result_Efield_{element}_" + f"{evaluation}.txt"[:10]}
        if element == "tri":
            names["iteration_error"] = f"iteration_error_{element}" + suffix
        return names
    nr = f"{int(index):03d}"
    if element == "quad":
        return {"result": f"Result_SEH_quad_{form}_{nr}.txt",
                "efield": f"Result_Efield_quad_{evaluation}_{nr}.txt"}
    if multi_mesh:
        return {"result": f"Result_SEH_tri_{evaluation}_{nr}.txt",
                "efield": f"Result_Efield_tri_{evaluation}_{nr}.txt",
                "iteration_error": f"iteration_error_{form}.txt" + suffix}
    return {"result": f"Result_SEH_tri_{form}_{nr}.txt",
            "efield": f"Result_Efield_tri_{evaluation}_{nr}.txt",
            "iteration_error": f"iteration_error_{form}_{nr}" + suffix}
