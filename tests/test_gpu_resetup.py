"""A handle set up again (the loop over surface realisations): a small mesh,
then a larger one with one more octree level, then the small one again.
Each time the matvec, a near-field preconditioned solve and a batched solve
on the reused handle equal those of a fresh handle on the same mesh (the
device buffers are reused where they fit; nothing of the previous mesh may
leak into the next operator).  A failed set-up leaves the handle unusable
until a set-up succeeds."""
import numpy as np
import pytest

from specklesim_amd import DeviceBuffer, SieTriMlfmm, meshes, plane_wave
from specklesim_amd._native import SpeckleError

pytestmark = pytest.mark.gpu
LAM = 633e-9
KW = dict(formulation="ICTF", normal_convention=2, precision="mixed")


def run(s, x):
    y = s.calculate_vector_b_tri(x)
    k, e = plane_wave(0.0, 0.0, 2, "TM")
    b = s.get_vector_b_tri(k, e)
    xs, info = s.gmres_run(b, tolerance=1e-6, max_iterations=650, use_initial_guess=False,
                           near_field=True)
    assert info["info"] == 0
    nb = 2 * s.N * 16
    bs, xb = [], []
    for th in np.linspace(0.0, 0.8, 9):
        k, e = plane_wave(th, 0.0, 2, "TM")
        bs.append(DeviceBuffer(nb).upload(s.get_vector_b_tri(k, e)))
        xb.append(DeviceBuffer(nb))
    infos = s.gmres_run_batch_device([d.ptr for d in bs], [d.ptr for d in xb], tolerance=1e-6,
                                     max_iterations=650, use_initial_guess=False, near_field=True)
    assert all(i["info"] == 0 for i in infos)
    return y, xs, np.stack([d.download() for d in xb])


def test_handle_set_up_again():
    small = meshes.rough_surface(20, 20, 2.5, 2.5, 0.1, 0.5, 4, LAM)
    large = meshes.rough_surface(50, 50, 5.0, 5.0, 0.1, 0.5, 5, LAM)
    h = SieTriMlfmm(LAM, 2.25, **KW)
    levels = []
    for mesh in (small, large, small):
        h.precalculation_tri(*mesh)
        levels.append(h.stats()["l_max"])
        fresh = SieTriMlfmm(LAM, 2.25, **KW).precalculation_tri(*mesh)
        rng = np.random.default_rng(h.N)
        x = rng.standard_normal(2 * h.N) + 1j * rng.standard_normal(2 * h.N)
        y1, s1, b1 = run(h, x)
        y2, s2, b2 = run(fresh, x)
        assert np.array_equal(y1, y2)
        assert np.array_equal(s1, s2)
        assert np.array_equal(b1, b2)
        fresh.close()
    assert levels[1] == levels[0] + 1


def test_failed_setup_leaves_handle_unusable():
    h = SieTriMlfmm(LAM, 2.25, **KW).precalculation_tri(
        *meshes.rough_surface(10, 10, 1.0, 1.0, 0.1, 0.5, 4, LAM))
    # two triangles that share only a node: no RWG edge
    pp = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, -1, 0]], float) * 1e-7
    tt = np.array([[1, 2, 3], [1, 4, 5]], np.int32)
    with pytest.raises(SpeckleError):
        h.precalculation_tri(pp, tt)
    with pytest.raises(SpeckleError, match="spk_setup_tri"):
        h.calculate_vector_b_tri(np.zeros(2 * max(h.N, 1), np.complex128))
    with pytest.raises(SpeckleError, match="spk_setup_tri"):
        h.get_vector_b_tri([0, 0, -1], [1, 0, 0])
