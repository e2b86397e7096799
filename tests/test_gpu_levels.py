"""Per-level truncation numbers (spk_set_level_truncation): SS with the
interpolation of the children's coefficients to the parent's grid
(Lagrange_Interpolation_New, src/lib_sie_mlfmm.f90:268-390, in
src/lib_sie_tri_mlfmm_interface.f90:466-528) and RR (:600-655) either as
the reference's code does it or by anterpolation, against the oracle's
restatement of the same operator:

- a 3 lambda patch with levels 2-3 and a 5 lambda patch with levels 2-4,
  K sized to the boxes (16/12 and 19/15/12): the matvec equals the
  oracle's to 1e-10 in both downward modes;
- truncation numbers equal to the default through the same call give the
  default operator bit for bit;
- a near-field preconditioned solve on such a handle reaches the
  tolerance in its true residual;
- mixed handles, the transposed operator and the batched products refuse
  such a handle.
"""
import numpy as np
import pytest

from oracle import oracle as O
from specklesim_amd import DeviceBuffer, SieTriMlfmm, meshes, plane_wave
from specklesim_amd._native import SpeckleError

pytestmark = pytest.mark.gpu
LAM = 633e-9

CASES = {
    "3 lambda, levels 2-3": (lambda: meshes.rough_surface(30, 30, 3.0, 3.0, 0.1, 0.5, 2, LAM),
                             60, [0, 0, 16, 12], (2, 3)),
    "5 lambda, levels 2-4": (lambda: meshes.rough_surface(50, 50, 5.0, 5.0, 0.1, 0.5, 3, LAM),
                             40, [0, 0, 19, 15, 12], (2, 4)),
}


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _gpu(pp, tt, s_opt, K=None, rr="anterpolate", precision="fp64"):
    s = SieTriMlfmm(LAM, 2.25, formulation="ICTF", normal_convention=2, tree_s_opt=s_opt,
                    precision=precision)
    if K is not None:
        s.set_level_truncation(K, rr)
    return s.precalculation_tri(pp, tt)


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("rr", ["anterpolate", "reference"])
def test_matvec_equals_the_oracle(name, rr):
    mesh, s_opt, K, levels = CASES[name]
    pp, tt = mesh()
    g = _gpu(pp, tt, s_opt, K, rr)
    st = g.stats()
    assert (st["l_min"], st["l_max"]) == levels
    o = O.Oracle(pp, tt, LAM, 2.25, formulation="ICTF", normal_convention=2, s_opt=s_opt)
    o.tree()
    o.near_csr(nthreads=16)
    o.mlfmm_setup(16, level_K=K, rr={"anterpolate": "adjoint", "reference": "reference"}[rr])
    rng = np.random.default_rng(11)
    x = rng.standard_normal(2 * g.N) + 1j * rng.standard_normal(2 * g.N)
    yo = o.matvec(x, 16)
    yg = g.calculate_vector_b_tri(x)
    e = rel(yg, yo)
    print(f"{name} {rr}: GPU vs oracle {e:.2e}")
    assert e < 1e-10


def test_default_values_give_the_default_operator():
    mesh, s_opt, _, _ = CASES["3 lambda, levels 2-3"]
    pp, tt = mesh()
    g0 = _gpu(pp, tt, s_opt)
    g1 = _gpu(pp, tt, s_opt, [19, 19, 19, 19])
    x = np.random.default_rng(2).standard_normal(2 * g0.N) + 0j
    assert np.array_equal(g0.calculate_vector_b_tri(x), g1.calculate_vector_b_tri(x))


def test_solve_and_refusals():
    mesh, s_opt, K, _ = CASES["5 lambda, levels 2-4"]
    pp, tt = mesh()
    g = _gpu(pp, tt, s_opt, K)
    k, e = plane_wave(0.0, 0.0, 2, "TM")
    nb = 2 * g.N * 16
    db, dx = DeviceBuffer(nb), DeviceBuffer(nb)
    g.get_vector_b_tri_device(k, e, db.ptr)
    info = g.gmres_run_device(db.ptr, dx.ptr, tolerance=1e-6, max_iterations=650,
                              orthogonalization="ICGS", use_initial_guess=False, near_field=True)
    assert info["info"] == 0, info
    x = dx.download()
    b = db.download()
    assert rel(g.calculate_vector_b_tri(x), b) <= 1.01e-6
    with pytest.raises(SpeckleError):
        g.set_transposed(True)
    with pytest.raises(SpeckleError):
        g.calculate_vector_b_tri_batch(np.zeros((2, 2 * g.N), complex))
    with pytest.raises(SpeckleError):
        _gpu(pp, tt, s_opt, K, precision="mixed")
    with pytest.raises(SpeckleError):
        _gpu(pp, tt, s_opt, [0, 0, 12, 16, 12])
