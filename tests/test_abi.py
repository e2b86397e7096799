"""The C-ABI library loads and exports every entry point include/*.h declares
(no GPU work is started here)."""
import ctypes
import glob
import os
import re

import numpy as np

from specklesim_amd import _native, meshes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(spk_\w+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = _native.lib()
    declared = _declared()
    assert len(declared) >= 19
    for name in declared:
        assert hasattr(lib, name), f"{name} declared in include/ but not exported"
    assert sorted(_native.EXPORTS) == declared


def test_last_error_is_a_string():
    assert isinstance(_native.lib().spk_last_error(), (bytes, type(None)))


def test_create_rejects_bad_arguments():
    p = _native.Params(633e-9, 1, 0, 2.25, 0, 99, 1, 3, 250, 19, 0, 0)
    h = ctypes.c_void_p()
    assert _native.lib().spk_create(ctypes.byref(p), ctypes.byref(h)) == -1
    assert b"formulation" in _native.lib().spk_last_error()


def test_topology_probe_runs_without_gpu():
    pp, tt = meshes.icosphere(1, 1.0)
    L = _native.lib()
    n = ctypes.c_int64()
    pp = np.ascontiguousarray(pp)
    tt = np.ascontiguousarray(tt, dtype=np.int32)
    rc = L.spk_topology_probe(pp.shape[0], pp.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                              tt.shape[0], tt.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), 1,
                              250, ctypes.byref(n), None, None, None, None, None, None, None)
    assert rc == 0 and n.value == 120  # 80 triangles -> 120 edges on a closed surface
