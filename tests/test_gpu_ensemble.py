"""spk_ensemble_intensity (no counterpart in the reference, whose angle loop
solves one realisation and one angle at a time): several realisations of
different sizes, several incidence angles each, solved by worker threads
with their own handles, reduced to scattered intensities.

- Each (realisation, angle) intensity equals the same solve made one handle
  at a time (the batched near-field preconditioned solve of that
  realisation's angles, then spk_scattered_field_device) to 1e-12: the
  ensemble changes the scheduling, not the arithmetic;
- one and three workers give the same intensities;
- against the oracle: the FP64 restatement's GMRES solution (IMGS, tolerance
  1e-9) of one realisation and angle, through the restatement's scattered
  field, agrees with the ensemble's intensity to 1e-4 (the far-field
  accuracy the ensemble's 1e-6 tolerance buys);
- a bad mesh is rejected before any device work, naming the mesh.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from specklesim_amd import DeviceBuffer, SieTriMlfmm, ensemble_intensity, meshes, plane_wave
from specklesim_amd import _native as nat

pytestmark = pytest.mark.gpu
LAM = 633e-9


def _case():
    ms = [meshes.rough_surface(c, c, c / 10, c / 10, 0.1, 0.5, seed, LAM)
          for c, seed in ((20, 1), (28, 2), (24, 3))]
    angles = [plane_wave(np.radians(th), 0.0, 2, "TM") for th in (0.0, 20.0, 40.0, 60.0, 10.0)]
    th = np.radians(np.linspace(-70, 70, 15))
    pts = 1e3 * LAM * np.stack([np.sin(th), np.zeros_like(th), np.cos(th)], axis=1)
    return ms, angles, pts


def test_ensemble_equals_one_handle_at_a_time():
    ms, angles, pts = _case()
    I3, info3 = ensemble_intensity(LAM, 2.25, ms, angles, pts, workers=3)
    I1, info1 = ensemble_intensity(LAM, 2.25, ms, angles, pts, workers=1)
    assert I3.shape == (3, 5, 15)
    assert all(i["info"] == 0 for row in info3 for i in row)
    assert np.max(np.abs(I3 - I1)) <= 1e-12 * np.max(np.abs(I1))
    for m, (pp, tt) in enumerate(ms):
        s = SieTriMlfmm(LAM, 2.25, formulation="ICTF", normal_convention=2,
                        precision="mixed").precalculation_tri(pp, tt)
        nb = 2 * s.N * 16
        bs = [DeviceBuffer(nb) for _ in angles]
        xs = [DeviceBuffer(nb) for _ in angles]
        for b, (k, e) in zip(bs, angles):
            s.get_vector_b_tri_device(k, e, b.ptr)
        infos = s.gmres_run_batch_device([b.ptr for b in bs], [x.ptr for x in xs],
                                         max_iterations=650, tolerance=1e-6,
                                         use_initial_guess=False, near_field=True)
        for a, x in enumerate(xs):
            E, _ = s.scattered_field_device(pts, x.ptr)
            Ia = np.sum(np.abs(E) ** 2, axis=1)
            assert np.max(np.abs(Ia - I3[m, a])) <= 1e-12 * np.max(Ia), (m, a)
            assert infos[a]["iters"] == info3[m][a]["iters"]


def test_ensemble_against_the_oracle():
    ms, angles, pts = _case()
    I, info = ensemble_intensity(LAM, 2.25, ms[:1], angles[:2], pts, workers=2)
    pp, tt = ms[0]
    o = O.Oracle(pp, tt, LAM, 2.25, formulation="ICTF", normal_convention=2)
    o.tree()
    o.near_csr(nthreads=16)
    o.mlfmm_setup(16)
    k, e = angles[1]
    xo, io = o.gmres(o.rhs_plane(k, e), tol=1e-9, maxit=650, nthreads=16)
    assert io["info"] == 0
    Io = np.sum(np.abs(o.scattered_field(pts, xo)[0]) ** 2, axis=1)
    err = np.linalg.norm(I[0, 1] - Io) / np.linalg.norm(Io)
    print(f"ensemble intensity vs the oracle's FP64 solve: {err:.2e}")
    assert err < 1e-4


def test_ensemble_rejects_a_bad_mesh():
    L = nat.lib()
    p = nat.Params(LAM, 1.0, 0.0, 2.25, 0.0, 3, 2, 3, 250, 19, 1, 0)
    pp = np.zeros((2, 3))
    tt = np.ones((1, 3), np.int32)
    ms = (nat.Mesh * 1)(nat.Mesh(2, pp.ctypes.data_as(C.POINTER(C.c_double)), 1,
                                 tt.ctypes.data_as(C.POINTER(C.c_int))))
    k = np.array([0.0, 0.0, -1.0])
    e = np.array([1.0, 0.0, 0.0])
    out = np.zeros(1)
    infos = (nat.GmresInfo * 1)()
    gp = nat.GmresParams(650, 0, 1e-6, 3, 0, 0, 0.0, 2, 0)
    dp = C.POINTER(C.c_double)
    rc = L.spk_ensemble_intensity(C.byref(p), 1, ms, 1, k.ctypes.data_as(dp), e.ctypes.data_as(dp),
                                  0, None, C.byref(gp), 1, out.ctypes.data_as(dp), infos)
    assert rc == -1 and b"mesh 0" in L.spk_last_error()
