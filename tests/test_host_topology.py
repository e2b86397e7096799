"""The product's host setup (RWG topology and octree, spk_topology_probe) is
bit-exact with the restatement; runs on the CPU."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from specklesim_amd import _native, meshes

LAM = 633e-9


def probe(pp, tt, conv=1, s_opt=250):
    L = _native.lib()
    pp = np.ascontiguousarray(pp, dtype=np.float64)
    tt = np.ascontiguousarray(tt, dtype=np.int32)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
    n = C.c_int64()
    _native.check(L.spk_topology_probe(pp.shape[0], dp(pp), tt.shape[0], ip(tt), conv, s_opt,
                                       C.byref(n), None, None, None, None, None, None, None))
    N = n.value
    corner = np.zeros((N, 2), np.int32)
    elem = np.zeros((N, 2), np.int32)
    leaf = np.zeros(N, np.int64)
    a, b = C.c_int(), C.c_int()
    D = np.zeros(1)
    xm = np.zeros(3)
    _native.check(L.spk_topology_probe(pp.shape[0], dp(pp), tt.shape[0], ip(tt), conv, s_opt,
                                       C.byref(n), ip(corner), ip(elem),
                                       leaf.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(a),
                                       C.byref(b), dp(D), dp(xm)))
    return dict(corner=corner, elem=elem, leaf=leaf, l_min=a.value, l_max=b.value, D=D[0], xmin=xm)


MESHES = {
    "icosphere3": lambda: meshes.icosphere(3, 0.5 * LAM),
    "octasphere258": lambda: meshes.octasphere(258, 0.5e-6),
    "rough20": lambda: meshes.rough_surface(20, 20, 2.0, 2.0, 0.1, 0.5, 4, LAM),
    "rough100": lambda: meshes.rough_surface(100, 100, 10.0, 10.0, 0.1, 0.5, 1, LAM),
}


@pytest.mark.parametrize("name", sorted(MESHES))
@pytest.mark.parametrize("s_opt", [250, 30])
def test_topology_and_tree_bit_exact(name, s_opt):
    pp, tt = MESHES[name]()
    got = probe(pp, tt, 1, s_opt)
    o = O.Oracle(pp, tt, LAM, 2.25, s_opt=s_opt)
    c, e = o.pairs()
    assert np.array_equal(got["corner"], c)
    assert np.array_equal(got["elem"], e)
    t = o.tree()
    assert got["l_min"] == t["l_min"] and got["l_max"] == t["l_max"]
    assert got["D"] == t["D"]
    assert np.array_equal(got["xmin"], t["xmin"])
    assert np.array_equal(got["leaf"], t["leaf"])


def test_config2_tree_shape():
    pp, tt = MESHES["rough100"]()
    got = probe(pp, tt, 2, 250)
    assert got["leaf"].size == 29800  # 100x100 cells: 30200 edges minus 400 boundary
    assert got["l_max"] == 4


def test_bad_mesh_is_rejected():
    pp, tt = meshes.icosphere(1, 1.0)
    tt = tt.copy()
    tt[0, 0] = 0  # 1-based indices required
    with pytest.raises(_native.SpeckleError):
        probe(pp, tt)


def test_single_edge_mesh():
    pp = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0.1]], float) * 1e-7
    tt = np.array([[1, 2, 3], [2, 4, 3]], np.int32)
    got = probe(pp, tt)
    assert got["leaf"].size == 1 and got["l_max"] == 1
