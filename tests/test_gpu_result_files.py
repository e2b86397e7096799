"""The run's output files from a GPU solve against the oracle's solve of the
same inputs (the mesh read back from pp_/tt_ files, as calc_p(5) = 4 reads
it): Result_SEH_* (n_el, m_pairs, then 2N lines of re/im in ES19.12,
src/lib_sie_solver_gmres_interface.f90:164-170) and iteration_error_* (the
convergence checks of the GMRES callback,
src/lib/libmath_external/src/algorithm_842/zPackgmres.f90:1337-1352)."""
import numpy as np
import pytest

from oracle import oracle as O
from specklesim_amd import SieTriMlfmm, meshes, plane_wave
from specklesim_amd import results as R

pytestmark = pytest.mark.gpu
LAM = 633e-9


def test_result_files_of_gpu_and_oracle_solves(tmp_path):
    pp, tt = meshes.rough_surface(24, 24, 2.4, 2.4, 0.1, 0.5, 6, LAM)
    meshes.write_pp_tt(pp, tt, tmp_path / "pp_patch.txt", tmp_path / "tt_patch.txt")
    pp, tt = meshes.read_pp_tt(tmp_path / "pp_patch.txt", tmp_path / "tt_patch.txt")
    g = SieTriMlfmm(LAM, 2.25, formulation="ICTF", normal_convention=2).precalculation_tri(pp, tt)
    o = O.Oracle(pp, tt, LAM, 2.25, formulation="ICTF", normal_convention=2)
    o.tree()
    o.near_csr(nthreads=16)
    o.mlfmm_setup(16)
    k, e = plane_wave(0.0, 0.0, 2, "TM")
    tol, maxit = 1e-6, 400
    xg, ig = g.gmres_run(g.get_vector_b_tri(k, e), tolerance=tol, max_iterations=maxit)
    xo, io = o.gmres(o.rhs_plane(k, e), tol=tol, maxit=maxit, nthreads=16)
    assert ig["info"] == io["info"] == 0
    names = R.output_names("ICTF")
    files = {}
    for who, x, hist in (("gpu", xg, ig["hist"]), ("oracle", xo, io["hist"])):
        d = tmp_path / who
        d.mkdir()
        R.write_result_seh(d / names["result"], tt.shape[0], x)
        R.write_iteration_error(d / names["iteration_error"], hist, tol, maxit)
        files[who] = (R.read_result_seh(d / names["result"]),
                      R.read_iteration_error(d / names["iteration_error"]))
    (ng, mg, yg), (itg, beg) = files["gpu"]
    (no, mo, yo), (ito, beo) = files["oracle"]
    assert (ng, mg) == (no, mo) == (tt.shape[0], g.N)
    assert yg.shape == yo.shape == (2 * g.N,)
    assert np.linalg.norm(yg - yo) / np.linalg.norm(yo) < 1e-6
    # the same convergence checks at the same steps, the same backward errors
    n = min(len(itg), len(ito))
    assert n >= 1 and abs(len(itg) - len(ito)) <= 1
    assert np.array_equal(itg[:n], ito[:n])
    assert np.max(np.abs(np.log(beg[:n] / beo[:n]))) < 1e-3
    # the file's 13 digits hold the solution: what was written is what was solved
    assert np.linalg.norm(yg - xg) <= 1e-12 * np.linalg.norm(xg)
