"""The batched matvec (R right-hand sides per launch) against R single
matvecs of the same handle.  The near-field and disaggregation sums run in
another order, and the far field amplifies rounding by the cancellation of
the K = 19 plane-wave sum at the 0.31-0.62 lambda leaves (about 1e4 here):
complex128 agrees to 1e-10 relative L2 per vector (1.3e-12 measured),
complex64 to 3e-3 against the FP64 operator, as the single complex64
operator does.  Complex64 batches of 9..16 run the matrix-core kernels
(near field, aggregation, disaggregation); smaller ones the vector kernels."""
import numpy as np
import pytest

from specklesim_amd import SieTriMlfmm, meshes

pytestmark = pytest.mark.gpu
LAM = 633e-9


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def solver():
    pp, tt = meshes.rough_surface(40, 40, 5.0, 5.0, 0.1, 0.5, 6, LAM)
    return SieTriMlfmm(LAM, 2.25, formulation="ICTF", normal_convention=2,
                       precision="mixed").precalculation_tri(pp, tt)


@pytest.mark.parametrize("R", [1, 3, 16])
def test_batched_matvec_fp64(solver, R):
    rng = np.random.default_rng(10 + R)
    X = rng.standard_normal((R, 2 * solver.N)) + 1j * rng.standard_normal((R, 2 * solver.N))
    Y = solver.calculate_vector_b_tri_batch(X)
    for r in range(R):
        assert rel(Y[r], solver.calculate_vector_b_tri(X[r])) < 1e-10, r


@pytest.mark.parametrize("R", [2, 11, 16])
def test_batched_matvec_c64(solver, R):
    rng = np.random.default_rng(20 + R)
    X = (rng.standard_normal((R, 2 * solver.N)) + 1j * rng.standard_normal((R, 2 * solver.N))
         ).astype(np.complex64)
    Y = solver.calculate_vector_b_tri_batch(X, c64=True)
    assert Y.dtype == np.complex64
    for r in range(R):
        ref = solver.calculate_vector_b_tri(X[r].astype(np.complex128))
        assert rel(Y[r].astype(np.complex128), ref) < 3e-3, r


def test_batched_matvec_rejects_bad_sizes(solver):
    from specklesim_amd._native import SpeckleError

    with pytest.raises(SpeckleError):
        solver.calculate_vector_b_tri_batch_device([0] * 17, [0] * 17)
    with pytest.raises(SpeckleError):
        solver.calculate_vector_b_tri_batch_device([], [])


def test_batched_mixed_solve(solver):
    """Four incidence angles solved together reach the FP64 tolerance in the
    FP64 residual, and each solution matches the single mixed solve of its
    right-hand side to the refinement accuracy (same algorithm, other
    summation order), and the FP64 solution's BRDF to 1e-4."""
    from specklesim_amd import DeviceBuffer, farfield, plane_wave

    angles = [0.0, 15.0, 30.0, 45.0]
    nb = 2 * solver.N * 16
    bs, xs, rhs = [], [], []
    for a in angles:
        k, e = plane_wave(np.deg2rad(a), 0.0, 2, "TM")
        b = solver.get_vector_b_tri(k, e)
        rhs.append(b)
        bs.append(DeviceBuffer(nb).upload(b))
        xs.append(DeviceBuffer(nb))
    infos = solver.gmres_run_batch_device([d.ptr for d in bs], [d.ptr for d in xs],
                                          max_iterations=1300, tolerance=1e-6,
                                          use_initial_guess=False)
    for r, (b, inf) in enumerate(zip(rhs, infos)):
        x = xs[r].download()
        assert inf["info"] == 0 and inf["be"] <= 1e-6, (r, inf)
        res = np.linalg.norm(b - solver.calculate_vector_b_tri(x)) / np.linalg.norm(b)
        assert abs(res - inf["be"]) < 1e-9 * max(1.0, res / 1e-6), r
        x1, i1 = solver.gmres_run(b, tolerance=1e-6, max_iterations=1300, use_initial_guess=False)
        assert rel(x, x1) < 1e-4, r
        _, I0 = farfield.brdf(solver, x1, np.deg2rad(angles[r]), (-1.2, 1.2), 41, 0.0, 1e3 * LAM)
        _, I1 = farfield.brdf(solver, x, np.deg2rad(angles[r]), (-1.2, 1.2), 41, 0.0, 1e3 * LAM)
        assert np.max(np.abs(I1 - I0)) / np.max(np.abs(I0)) < 1e-4, r


def test_batched_mixed_solve_wide(solver):
    """Twelve incidence angles (the matrix-core batch width) reach the FP64
    tolerance, each in its own FP64 residual."""
    from specklesim_amd import DeviceBuffer, plane_wave

    angles = np.linspace(0.0, 55.0, 12)
    nb = 2 * solver.N * 16
    bs, xs, rhs = [], [], []
    for a in angles:
        k, e = plane_wave(np.deg2rad(a), 0.0, 2, "TM")
        b = solver.get_vector_b_tri(k, e)
        rhs.append(b)
        bs.append(DeviceBuffer(nb).upload(b))
        xs.append(DeviceBuffer(nb))
    infos = solver.gmres_run_batch_device([d.ptr for d in bs], [d.ptr for d in xs],
                                          max_iterations=1300, tolerance=1e-6,
                                          use_initial_guess=False)
    for r, (b, inf) in enumerate(zip(rhs, infos)):
        assert inf["info"] == 0 and inf["be"] <= 1e-6, (r, inf)
        res = np.linalg.norm(b - solver.calculate_vector_b_tri(xs[r].download())) / np.linalg.norm(b)
        assert res <= 1e-6 and abs(res - inf["be"]) < 1e-9, (r, res, inf["be"])
