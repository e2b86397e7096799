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


def test_two_hip_runtimes_are_refused():
    """A process that maps two different libamdhip64 files (a PyTorch wheel's
    and ROCm's) is refused; two mappings of one file are one runtime."""
    one = ("7f0000000000-7f0000100000 r-xp 00000000 fd:01 1234 /opt/rocm/lib/libamdhip64.so.7.2.0\n"
           "7f0000100000-7f0000200000 r--p 00100000 fd:01 1234 /opt/rocm/lib/libamdhip64.so.7.2.0\n"
           "7f0000300000-7f0000400000 r-xp 00000000 fd:01 99 /usr/lib/libc.so.6\n")
    two = one + ("7f0000500000-7f0000600000 r-xp 00000000 fd:02 777 "
                 "/usr/local/lib/python3.10/dist-packages/torch/lib/libamdhip64.so\n")
    assert _native.hip_runtimes(one) == ["/opt/rocm/lib/libamdhip64.so.7.2.0"]
    assert len(_native.hip_runtimes(two)) == 2
    _native.check_single_hip_runtime(one)
    import pytest

    with pytest.raises(_native.SpeckleError, match="two HIP runtimes"):
        _native.check_single_hip_runtime(two)
    # this process, after loading the library
    _native.lib()
    with open("/proc/self/maps") as f:
        assert len(_native.hip_runtimes(f.read())) <= 1
