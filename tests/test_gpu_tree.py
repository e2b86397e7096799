"""The octree and its lists built on the GPU (spk_tree.hip) against the
host builder (SPK_HOST_TREE=1), which the oracle pins bit for bit
(tests/test_gpu_parity.py, tests/test_host_topology.py).

On rough patches of several sizes (levels 2-2 to 2-4), an icosphere (all
three axes populated), a patch small enough that l_max = 1, and a patch
with a lower s_opt: the leaf Morton indices, l_min, l_max, D and xmin, the
I4 list of every box at every level, the near-field CSR (ia, ja) and the
matvec are identical.
"""
import os

import numpy as np
import pytest

from specklesim_amd import SieTriMlfmm, meshes

pytestmark = pytest.mark.gpu
LAM = 633e-9

CASES = [
    ("patch 5 lambda", lambda: meshes.rough_surface(40, 40, 5.0, 5.0, 0.1, 0.5, 6, LAM), 250),
    ("patch 10 lambda", lambda: meshes.rough_surface(100, 100, 10.0, 10.0, 0.1, 0.5, 1, LAM), 250),
    ("patch 3 lambda, s_opt 60", lambda: meshes.rough_surface(30, 30, 3.0, 3.0, 0.1, 0.5, 4, LAM), 60),
    ("patch 1 lambda (l_max 1)", lambda: meshes.rough_surface(6, 6, 1.0, 1.0, 0.1, 0.5, 2, LAM), 250),
    ("icosphere", lambda: meshes.icosphere(3, 0.5 * LAM), 120),
]


def _build(pp, tt, s_opt, host):
    old = os.environ.get("SPK_HOST_TREE")
    os.environ["SPK_HOST_TREE"] = "1" if host else "0"
    try:
        return SieTriMlfmm(LAM, 2.25, formulation="ICTF", normal_convention=2,
                           tree_s_opt=s_opt).precalculation_tri(pp, tt)
    finally:
        if old is None:
            del os.environ["SPK_HOST_TREE"]
        else:
            os.environ["SPK_HOST_TREE"] = old


@pytest.mark.parametrize("name,mesh,s_opt", CASES, ids=[c[0] for c in CASES])
def test_device_tree_equals_host_tree(name, mesh, s_opt):
    pp, tt = mesh()
    gd = _build(pp, tt, s_opt, host=False)
    gh = _build(pp, tt, s_opt, host=True)
    td, th = gd.tree(), gh.tree()
    print(f"{name}: N {gd.N}, levels {td['l_min']}-{td['l_max']}")
    assert (td["l_min"], td["l_max"]) == (th["l_min"], th["l_max"])
    assert td["D"] == th["D"] and np.array_equal(td["xmin"], th["xmin"])
    assert np.array_equal(td["leaf"], th["leaf"])
    ia_d, ja_d, _ = gd.near_csr(values=False)
    ia_h, ja_h, _ = gh.near_csr(values=False)
    assert np.array_equal(ia_d, ia_h) and np.array_equal(ja_d, ja_h)
    lmax = td["l_max"]
    for level in range(td["l_min"], lmax + 1):
        for box in np.unique(td["leaf"] >> (3 * (lmax - level))):
            assert np.array_equal(gd.i4(level, int(box)), gh.i4(level, int(box))), (level, box)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(2 * gd.N) + 1j * rng.standard_normal(2 * gd.N)
    assert np.array_equal(gd.calculate_vector_b_tri(x), gh.calculate_vector_b_tri(x))
