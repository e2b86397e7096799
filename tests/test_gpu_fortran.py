"""The Fortran host binding (specklesim_amd/fortran/spk_gpu_binding.f90):
a Fortran program drives precalculation, matvec, right-hand side and GMRES
through ISO_C_BINDING and must reproduce the Python-side calls of the same
library on the same mesh."""
import os
import subprocess

import numpy as np
import pytest

from specklesim_amd import SieTriMlfmm, meshes

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "..", "specklesim_amd", "spk_fortran_check")
LAM = 633e-9


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built (no flang)")
def test_fortran_host_matches_python_host(tmp_path):
    pp, tt = meshes.rough_surface(16, 16, 2.0, 2.0, 0.1, 0.5, 4, LAM)
    fpp, ftt, fout = tmp_path / "pp.txt", tmp_path / "tt.txt", tmp_path / "out.txt"
    meshes.write_pp_tt(pp, tt, str(fpp), str(ftt))
    p = subprocess.run([EXE, str(fpp), str(ftt), str(fout)], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    with open(fout) as f:
        head = f.readline().split()
        data = np.loadtxt(f)
    n, its, be = int(head[0]), int(head[1]), float(head[2])
    y_f = data[: 2 * n, 0] + 1j * data[: 2 * n, 1]
    x_f = data[2 * n:, 0] + 1j * data[2 * n:, 1]

    # the same calls from Python (the driver's parameters: PMCHWT,
    # calc_p(7) = 2, lambda 633 nm, eps_r2 2.25)
    pp2, tt2 = meshes.read_pp_tt(str(fpp), str(ftt))
    g = SieTriMlfmm(LAM, 2.25, formulation="PMCHWT", normal_convention=2).precalculation_tri(pp2, tt2)
    assert g.N == n
    x = 1.0 + 1j * np.arange(1, 2 * n + 1) / (2 * n)
    y = g.calculate_vector_b_tri(x)
    assert np.linalg.norm(y_f - y) <= 1e-15 * np.linalg.norm(y)
    b = g.get_vector_b_tri([0.0, 0.0, -1.0], [1.0, 0.0, 0.0])
    xs, info = g.gmres_run(b, tolerance=1e-6, max_iterations=300, orthogonalization="IMGS")
    assert info["iters"] == its
    assert abs(info["be"] - be) <= 1e-14 * max(be, 1e-300)
    assert np.linalg.norm(x_f - xs) <= 1e-15 * np.linalg.norm(xs)
