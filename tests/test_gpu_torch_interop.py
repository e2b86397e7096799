"""torch tensors as device buffers of the *_device entry points.

torch ships its own HIP runtime.  Both libraries' HIP runtimes carry the
soname libamdhip64.so.7, so when torch is imported first, libspeckle_gpu.so
binds to the runtime already loaded and both share one HIP/HSA runtime.
This runs in a fresh interpreter so that the import order is the one
stated.
"""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r"""
import sys
import numpy as np
import torch
assert torch.cuda.is_available()
torch.zeros(1, device="cuda")
sys.path.insert(0, ROOT)
from specklesim_amd import SieTriMlfmm, meshes, plane_wave
lam = 633e-9
pp, tt = meshes.rough_surface(16, 16, 2.0, 2.0, 0.1, 0.5, 4, lam)
g = SieTriMlfmm(lam, 2.25, formulation="ICTF", normal_convention=2).precalculation_tri(pp, tt)
rng = np.random.default_rng(2)
x = rng.standard_normal(2 * g.N) + 1j * rng.standard_normal(2 * g.N)
dx = torch.from_numpy(x).to("cuda")
dy = torch.empty_like(dx)
torch.cuda.synchronize()
g.calculate_vector_b_tri_device(dx.data_ptr(), dy.data_ptr())
y = g.calculate_vector_b_tri(x)
assert np.array_equal(dy.cpu().numpy(), y)
k, e = plane_wave(0.0, 0.0, 2, "TM")
db = torch.empty_like(dx)
g.get_vector_b_tri_device(k, e, db.data_ptr())
assert np.array_equal(db.cpu().numpy(), g.get_vector_b_tri(k, e))
print("interop ok", flush=True)
""".replace("ROOT", repr(ROOT))


def test_torch_buffers_when_torch_is_imported_first():
    p = subprocess.run([sys.executable, "-c", SCRIPT], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-4000:]
    assert "interop ok" in p.stdout
