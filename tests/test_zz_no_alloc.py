"""spk_scattered_field_device_noalloc, the scattered-field entry point that
allocates nothing (no counterpart in the reference; for callers that cannot
allocate device memory), called through the C-ABI against
spk_scattered_field_device, the allocating entry point it mirrors.  Both run
the same kernel on the same inputs, so the fields must be bit-identical: on
two calls in a row, the second with fewer points (the staging array keeps
the first call's contents), and the call beyond the staging maximum is
rejected."""
import ctypes as C

import numpy as np
import pytest

from specklesim_amd import DeviceBuffer, SieTriMlfmm, meshes
from specklesim_amd import _native as nat

pytestmark = pytest.mark.gpu
LAM = 633e-9
MAX_POINTS = 16384
SPK_ERR_ARG = -1


def _points(n, seed):
    """n observation points on the upper hemisphere of radius 1000 lambda."""
    rng = np.random.default_rng(seed)
    th = np.arccos(rng.uniform(0.05, 1.0, n))
    ph = rng.uniform(0.0, 2 * np.pi, n)
    r = 1000 * LAM
    return np.ascontiguousarray(np.stack([r * np.sin(th) * np.cos(ph), r * np.sin(th) * np.sin(ph),
                                          r * np.cos(th)], axis=1))


def _call(fn, h, pts, d_x):
    E = np.zeros((pts.shape[0], 3), np.complex128)
    H = np.zeros((pts.shape[0], 3), np.complex128)
    dp = C.POINTER(C.c_double)
    rc = fn(h, pts.shape[0], pts.ctypes.data_as(dp), C.c_void_p(int(d_x)),
            E.view(np.float64).ctypes.data_as(dp), H.view(np.float64).ctypes.data_as(dp))
    return rc, E, H


def test_noalloc_equals_allocating_entry_point():
    L = nat.lib()
    pp, tt = meshes.rough_surface(20, 20, 2.0, 2.0, 0.1, 0.5, 4, LAM)
    s = SieTriMlfmm(LAM, 2.25, formulation="ICTF", normal_convention=2).precalculation_tri(pp, tt)
    rng = np.random.default_rng(8)
    x = rng.standard_normal(2 * s.N) + 1j * rng.standard_normal(2 * s.N)
    d_x = DeviceBuffer(2 * s.N * 16).upload(x)
    for npts, seed in ((3000, 1), (700, 2)):
        pts = _points(npts, seed)
        rc, E1, H1 = _call(L.spk_scattered_field_device_noalloc, s._h, pts, d_x.ptr)
        assert rc == 0, L.spk_last_error()
        rc, E0, H0 = _call(L.spk_scattered_field_device, s._h, pts, d_x.ptr)
        assert rc == 0, L.spk_last_error()
        assert np.abs(E0).max() > 0 and np.abs(H0).max() > 0
        assert np.array_equal(E1, E0) and np.array_equal(H1, H0), npts
    # zero points: nothing to do
    rc, _, _ = _call(L.spk_scattered_field_device_noalloc, s._h, np.zeros((0, 3)), d_x.ptr)
    assert rc == 0
    # beyond the staging maximum: rejected before any device work
    rc, _, _ = _call(L.spk_scattered_field_device_noalloc, s._h, np.zeros((MAX_POINTS + 1, 3)),
                     d_x.ptr)
    assert rc == SPK_ERR_ARG
    assert b"maximum" in L.spk_last_error()
