"""The adjoint gradient of a far-field intensity with respect to node
heights (spk_adjoint_height_gradient; no counterpart in the reference):
J = sum of |E|^2 over a 5 x 5 detector window of +-5 degrees around the
specular direction, a plane wave at 20 degrees on a rough patch.

- against central finite differences of J (step 1e-4 lambda, the forward
  solve to 1e-11 at z +- h) at 8 seeded nodes: to 1e-3 of the largest
  finite-difference value, on FP64 and mixed handles;
- the mixed handle's gradient equals the FP64 handle's to 1e-6 (both solve
  to an FP64 true residual of 1e-11);
- the torch.autograd.Function around it: backward gives the adjoint
  gradient (fresh interpreter, torch imported first).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from specklesim_amd import SieTriMlfmm, meshes, plane_wave
from specklesim_amd.adjoint import HeightIntensity, detector_window

pytestmark = pytest.mark.gpu
LAM = 633e-9
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _problem(precision, form="ICTF"):
    pp, tt = meshes.rough_surface(16, 16, 1.6, 1.6, 0.1, 0.5, 11, LAM)
    s = SieTriMlfmm(LAM, 2.25, formulation=form, normal_convention=2, precision=precision)
    th = np.radians(20.0)
    k, e = plane_wave(th, 0.0, 2, "TM")
    pts = detector_window(th, 5.0, 5, 1e3 * LAM)
    return HeightIntensity(s, pp, tt, k, e, pts, tolerance=1e-11, dz=1e-6 * LAM), pp


def _nodes(pp, count=8, seed=5):
    # interior nodes (the patch edge nodes carry no RWG of their own)
    x, y = pp[:, 0], pp[:, 1]
    inner = np.nonzero((np.abs(x) < 0.99 * np.abs(x).max()) & (np.abs(y) < 0.99 * np.abs(y).max()))[0]
    return np.sort(np.random.default_rng(seed).choice(inner, count, replace=False))


@pytest.fixture(scope="module")
def fd_reference():
    prob, pp = _problem("fp64")
    z0 = pp[:, 2].copy()
    nodes = _nodes(pp)
    h = 1e-4 * LAM
    g = []
    for i in nodes:
        zp, zm = z0.copy(), z0.copy()
        zp[i] += h
        zm[i] -= h
        g.append((prob.value(zp) - prob.value(zm)) / (2 * h))
    return z0, nodes, np.array(g), prob.value(z0)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_adjoint_gradient_matches_central_differences(fd_reference, precision):
    z0, nodes, gfd, J0 = fd_reference
    prob, _ = _problem(precision)
    J, g, info = prob.gradient(z0, nodes)
    assert info[0]["info"] == 0 and info[1]["info"] == 0, info
    assert abs(J - J0) <= 1e-9 * J0
    err = np.max(np.abs(g - gfd)) / np.max(np.abs(gfd))
    print(f"{precision}: adjoint vs central differences, max error {err:.2e} of max |dJ/dz| "
          f"{np.max(np.abs(gfd)):.3e}")
    assert err < 1e-3
    # the gradient is not trivially small anywhere it is compared
    assert np.min(np.abs(gfd)) > 1e-3 * np.max(np.abs(gfd))


def test_mixed_gradient_equals_fp64(fd_reference):
    z0, nodes, _, _ = fd_reference
    _, g64, _ = _problem("fp64")[0].gradient(z0, nodes)
    _, g32, _ = _problem("mixed")[0].gradient(z0, nodes)
    assert np.max(np.abs(g32 - g64)) <= 1e-6 * np.max(np.abs(g64))


SCRIPT = r"""
import sys
import numpy as np
import torch
sys.path.insert(0, ROOT)
from specklesim_amd import SieTriMlfmm, meshes, plane_wave
from specklesim_amd.adjoint import HeightIntensity, detector_window, torch_function
lam = 633e-9
pp, tt = meshes.rough_surface(12, 12, 1.2, 1.2, 0.1, 0.5, 3, lam)
s = SieTriMlfmm(lam, 2.25, formulation="ICTF", normal_convention=2)
k, e = plane_wave(0.0, 0.0, 2, "TM")
prob = HeightIntensity(s, pp, tt, k, e, detector_window(0.0, 5.0, 3, 1e3 * lam), tolerance=1e-11,
                       dz=1e-6 * lam)
z = torch.tensor(pp[:, 2], dtype=torch.float64, requires_grad=True)
J = torch_function(prob)(z)
J.backward()
J2, g, _ = prob.gradient(pp[:, 2])
assert abs(J.item() - J2) <= 1e-12 * J2
assert np.array_equal(z.grad.numpy(), g)
print("torch ok", J.item())
"""


def test_torch_function_backward_is_the_adjoint():
    p = subprocess.run([sys.executable, "-c", SCRIPT.replace("ROOT", repr(ROOT))],
                       capture_output=True, text=True, timeout=240)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "torch ok" in p.stdout


def test_adjoint_gradient_against_the_oracle(fd_reference):
    """Central differences of J computed entirely by the oracle (its MLFMM
    operator, CERFACS GMRES with IMGS to 1e-11 and its scattered field) at
    two of the nodes, against the adjoint gradient."""
    from oracle import oracle as O

    z0, nodes, gfd, _ = fd_reference
    prob, pp = _problem("fp64")
    _, g, _ = prob.gradient(z0, nodes[:2])
    h = 1e-4 * LAM

    def J_oracle(z):
        p = pp.copy()
        p[:, 2] = z
        o = O.Oracle(p, prob.tt, LAM, 2.25, formulation="ICTF", normal_convention=2)
        o.tree()
        o.near_csr(nthreads=16)
        o.mlfmm_setup(16)
        x, info = o.gmres(o.rhs_plane(prob.k_in, prob.e_in), tol=1e-11, maxit=650, nthreads=16)
        assert info["info"] == 0
        E = o.scattered_field(prob.points, x)[0]
        return float(np.sum(np.abs(E) ** 2))

    for i, gi in zip(nodes[:2], g):
        zp, zm = z0.copy(), z0.copy()
        zp[i] += h
        zm[i] -= h
        go = (J_oracle(zp) - J_oracle(zm)) / (2 * h)
        print(f"node {i}: adjoint {gi:.6e}, oracle central difference {go:.6e}")
        assert abs(gi - go) <= 1e-3 * np.max(np.abs(gfd))
