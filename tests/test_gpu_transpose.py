"""The transposed operator Z^T (spk_set_transposed; no counterpart in the
reference), the matrix of the adjoint system Z^T lambda = dJ/dI.

The checks use the bilinear identity u^T (Z v) = v^T (Z^T u) on seeded
random vectors, for the near blocks and the MLFMM far part separately
(spk_matvec_parts), so that a wrong block, stage or coefficient anywhere
in the transposed pipeline shows: measured 2e-17 (near) and at most
5.5e-14 (far) of |u| |Z v|.

The mesh is a 5 lambda patch (40 x 40 cells, N = 4720, octree levels 2-3:
SS, M2L on two levels, RR).  Smaller leaf boxes than the tree's default
(e.g. a 2 lambda patch split to 0.25 lambda leaves) are avoided: there
the K = 19 translation sum cancels terms many orders of magnitude larger
than its result, so any reordering of the far-field arithmetic, the
transposed pipeline's included, moves the product by about 1e-5 in FP64
(and the complex64 operator is useless on such boxes, forward or
transposed).

- every formulation class (PMCHWT-like, ICTF's row scalings, MCTF2's
  coefficients) to 1e-12, and switching back restores Z bit for bit;
- against the oracle's MLFMM operator (u^T Z_oracle v against
  v^T Z^T u) to 1e-10;
- the complex64 operator of a mixed handle transposes too (to its 3e-3
  bound against FP64; 6.7e-4 measured), and a near-field preconditioned
  solve of Z^T lambda = r reaches the tolerance in its true residual under
  the FP64 transposed operator.
"""
import numpy as np
import pytest

from oracle import oracle as O
from specklesim_amd import DeviceBuffer, SieTriMlfmm, meshes, plane_wave

pytestmark = pytest.mark.gpu
LAM = 633e-9


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def patch():
    return meshes.rough_surface(40, 40, 5.0, 5.0, 0.1, 0.5, 6, LAM)


def _pairs(n, seed, count=3):
    rng = np.random.default_rng(seed)
    U = rng.standard_normal((count, n)) + 1j * rng.standard_normal((count, n))
    V = rng.standard_normal((count, n)) + 1j * rng.standard_normal((count, n))
    return U, V


@pytest.mark.parametrize("form", ["PMCHWT", "ICTF", "MCTF2"])
def test_transpose_bilinear_identity(patch, form):
    pp, tt = patch
    s = SieTriMlfmm(LAM, 2.25, formulation=form, normal_convention=2).precalculation_tri(pp, tt)
    st = s.stats()
    assert (st["l_min"], st["l_max"]) == (2, 3)
    n = 2 * s.N
    U, V = _pairs(n, 3)
    ZV = [s.matvec_parts(v) for v in V]
    y0 = s.calculate_vector_b_tri(V[0])
    s.set_transposed(True)
    assert s.transposed
    ZTU = [s.matvec_parts(u) for u in U]
    for i in range(len(U)):
        for part, tol in ((0, 1e-14), (1, 1e-12)):
            a, b = U[i] @ ZV[i][part], V[i] @ ZTU[i][part]
            err = abs(a - b) / (np.linalg.norm(U[i]) * np.linalg.norm(ZV[i][part]))
            assert err < tol, (form, i, part, err)
    # the operator really is not symmetric (so the identity has teeth)
    yt = s.calculate_vector_b_tri(V[0])
    assert rel(yt, y0) > 1e-2
    s.set_transposed(False)
    assert not s.transposed
    assert np.array_equal(s.calculate_vector_b_tri(V[0]), y0)


def test_transpose_against_the_oracle(patch):
    pp, tt = patch
    o = O.Oracle(pp, tt, LAM, 2.25, formulation="ICTF", normal_convention=2)
    o.tree()
    o.near_csr(nthreads=16)
    o.mlfmm_setup(16)
    s = SieTriMlfmm(LAM, 2.25, formulation="ICTF",
                    normal_convention=2).precalculation_tri(pp, tt).set_transposed(True)
    U, V = _pairs(2 * o.N, 4, 2)
    for u, v in zip(U, V):
        zv = o.matvec(v, 16)
        a, b = u @ zv, v @ s.calculate_vector_b_tri(u)
        assert abs(a - b) / (np.linalg.norm(u) * np.linalg.norm(zv)) < 1e-10


def test_transpose_complex64_and_solve(patch):
    pp, tt = patch
    s = SieTriMlfmm(LAM, 2.25, formulation="ICTF", normal_convention=2,
                    precision="mixed").precalculation_tri(pp, tt)
    s64 = SieTriMlfmm(LAM, 2.25, formulation="ICTF", normal_convention=2).precalculation_tri(pp, tt)
    k, e = plane_wave(0.3, 0.0, 2, "TM")
    r = np.conj(s64.get_vector_b_tri(k, e))
    s.set_transposed(True)
    s64.set_transposed(True)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(2 * s.N) + 1j * rng.standard_normal(2 * s.N)
    yt = s64.calculate_vector_b_tri(x)
    y32 = s.calculate_vector_b_tri_c64(x.astype(np.complex64)).astype(np.complex128)
    assert rel(y32, yt) < 3e-3
    # adjoint-style solve: Z^T lambda = r
    nb = 2 * s.N * 16
    db, dx = DeviceBuffer(nb).upload(r), DeviceBuffer(nb)
    info = s.gmres_run_device(db.ptr, dx.ptr, tolerance=1e-8, max_iterations=650,
                              orthogonalization="ICGS", use_initial_guess=False, near_field=True)
    assert info["info"] == 0, info
    lam = dx.download()
    res = rel(s64.calculate_vector_b_tri(lam), r)
    assert res <= 1.01e-8, res
