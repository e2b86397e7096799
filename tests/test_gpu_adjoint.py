"""The adjoint gradient of a far-field intensity with respect to node
heights (spk_adjoint_height_gradient; no counterpart in the reference):
J = sum of |E|^2 over a 5 x 5 detector window of +-5 degrees around the
specular direction, a plane wave at 20 degrees on a 2.4 lambda rough patch
(24 x 24 cells, N = 1680, leaf level 2 with 156 I4 pairs, so the far
field's share is exercised).

Finite-difference steps: the near blocks switch between the singular and
the regular quadrature where a triangle pair's centroid distance crosses
1.5 r_ref, and J jumps there.  On this family of meshes a step of 1e-4
lambda already crosses such a switch at some nodes (on a 1.6 lambda patch
node 218's -h mesh has 3 fewer singular pairs; the central difference reads
-4.31 against -0.1357 at steps of 1e-5 .. 1e-7 lambda), so every difference
here uses 1e-5 lambda and asserts that both stencil meshes keep the
unmoved mesh's singular-pair count.

- the gradient against central differences of J (forward solves to 1e-11)
  at 8 seeded nodes, to 1e-3 of the largest value (6.4e-5 measured), on
  FP64 and mixed handles; the mixed handle's gradient equals the FP64
  handle's to 1e-6 (4.9e-11 measured);
- each share (detector field, right-hand side, near blocks, far field)
  against its own brute-force central difference, computed with forward
  products at the moved meshes and the currents I and adjoint vector nu of
  separate GPU solves, to 1e-4 of the largest share (1.7e-6 measured);
- the same shares from the oracle's operator, right-hand side and field
  at the moved meshes (I and nu as above), their sum to 1e-4 (equal to the
  7 printed digits);
- the torch.autograd.Function around it: backward gives the adjoint
  gradient (fresh interpreter, torch imported first).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from specklesim_amd import DeviceBuffer, SieTriMlfmm, meshes, plane_wave
from specklesim_amd.adjoint import HeightIntensity, detector_window

pytestmark = pytest.mark.gpu
LAM = 633e-9
H_FD = 1e-5 * LAM
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TH = np.radians(20.0)


def _mesh():
    return meshes.rough_surface(24, 24, 2.4, 2.4, 0.1, 0.5, 11, LAM)


def _problem(precision, form="ICTF"):
    pp, tt = _mesh()
    s = SieTriMlfmm(LAM, 2.25, formulation=form, normal_convention=2, precision=precision)
    k, e = plane_wave(TH, 0.0, 2, "TM")
    pts = detector_window(TH, 5.0, 5, 1e3 * LAM)
    return HeightIntensity(s, pp, tt, k, e, pts, tolerance=1e-11, dz=1e-6 * LAM), pp


def _nodes(pp, count=8, seed=5):
    # interior nodes (the patch edge nodes carry no RWG of their own)
    x, y = pp[:, 0], pp[:, 1]
    inner = np.nonzero((np.abs(x) < 0.99 * np.abs(x).max()) & (np.abs(y) < 0.99 * np.abs(y).max()))[0]
    return np.sort(np.random.default_rng(seed).choice(inner, count, replace=False))


def _handle(pp, tt, z):
    p = pp.copy()
    p[:, 2] = z
    return SieTriMlfmm(LAM, 2.25, formulation="ICTF", normal_convention=2).precalculation_tri(p, tt)


@pytest.fixture(scope="module")
def fd_reference():
    prob, pp = _problem("fp64")
    z0 = pp[:, 2].copy()
    nodes = _nodes(pp)
    J0 = prob.value(z0)
    sing0 = prob.s.stats()["singular_pairs"]
    g = []
    for i in nodes:
        zp, zm = z0.copy(), z0.copy()
        zp[i] += H_FD
        zm[i] -= H_FD
        Jp = prob.value(zp)
        assert prob.s.stats()["singular_pairs"] == sing0, ("+h crosses a branch switch", i)
        Jm = prob.value(zm)
        assert prob.s.stats()["singular_pairs"] == sing0, ("-h crosses a branch switch", i)
        g.append((Jp - Jm) / (2 * H_FD))
    return z0, nodes, np.array(g), J0


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
    err = np.max(np.abs(g32 - g64)) / np.max(np.abs(g64))
    print(f"mixed vs fp64 gradient: {err:.2e}")
    assert err <= 1e-6


@pytest.fixture(scope="module")
def adjoint_state():
    """I (Z I = b) and nu (Z^T nu = F^T W conj(E)) from separate GPU solves,
    F's columns from the field of unit currents, and the adjoint's shares at
    three nodes."""
    pp, tt = _mesh()
    z0 = pp[:, 2].copy()
    s = _handle(pp, tt, z0)
    k, e = plane_wave(TH, 0.0, 2, "TM")
    pts = detector_window(TH, 5.0, 5, 1e3 * LAM)
    w = np.ones(len(pts))
    n = 2 * s.N

    def solve(rhs):
        nb = n * 16
        db, dx = DeviceBuffer(nb).upload(rhs), DeviceBuffer(nb)
        info = s.gmres_run_device(db.ptr, dx.ptr, tolerance=1e-11, max_iterations=650,
                                  orthogonalization="ICGS", use_initial_guess=False,
                                  near_field=True)
        assert info["info"] == 0, info
        return dx.download()

    st = s.stats()
    assert st["i4_pairs_per_level"][st["l_max"]] > 0
    I = solve(s.get_vector_b_tri(k, e))
    E = s.scattered_field(pts, I)[0].reshape(-1)
    F = np.zeros((E.size, n), complex)
    u = np.zeros(n, complex)
    for j in range(n):
        u[:] = 0
        u[j] = 1
        F[:, j] = s.scattered_field(pts, u)[0].reshape(-1)
    s.set_transposed(True)
    nu = solve(F.T @ (np.repeat(w, 3) * np.conj(E)))
    s.set_transposed(False)
    nodes = _nodes(pp, 3)
    _, g, _, terms = s.adjoint_height_gradient(k, e, pts, w, nodes, dz=1e-6 * LAM,
                                               tolerance=1e-11, return_terms=True)
    return dict(pp=pp, tt=tt, z0=z0, k=k, e=e, pts=pts, w=np.repeat(w, 3), I=I, E=E, nu=nu,
                nodes=nodes, g=g, terms=terms, sing=st["singular_pairs"])


def test_adjoint_shares_match_brute_force(adjoint_state):
    a = adjoint_state
    assert np.allclose(a["terms"].sum(axis=1), a["g"], rtol=1e-12, atol=0)
    for i, nd in enumerate(a["nodes"]):
        zp, zm = a["z0"].copy(), a["z0"].copy()
        zp[nd] += H_FD
        zm[nd] -= H_FD
        sp, sm = _handle(a["pp"], a["tt"], zp), _handle(a["pp"], a["tt"], zm)
        assert sp.stats()["singular_pairs"] == sm.stats()["singular_pairs"] == a["sing"]
        dE = sp.scattered_field(a["pts"], a["I"])[0].reshape(-1) - \
            sm.scattered_field(a["pts"], a["I"])[0].reshape(-1)
        fld = 2 * np.real(np.conj(a["E"]) * a["w"] @ dE) / (2 * H_FD)
        rhs = 2 * np.real(a["nu"] @ (sp.get_vector_b_tri(a["k"], a["e"]) -
                                     sm.get_vector_b_tri(a["k"], a["e"]))) / (2 * H_FD)
        (np_, fp), (nm, fm) = sp.matvec_parts(a["I"]), sm.matvec_parts(a["I"])
        near = -2 * np.real(a["nu"] @ (np_ - nm)) / (2 * H_FD)
        far = -2 * np.real(a["nu"] @ (fp - fm)) / (2 * H_FD)
        brute = np.array([fld, rhs, near, far])
        err = np.max(np.abs(a["terms"][i] - brute)) / np.max(np.abs(brute))
        print(f"node {nd}: shares {a['terms'][i]}, brute force {brute}, max error {err:.1e}")
        assert abs(far) > 1e-2 * np.max(np.abs(brute))
        assert err < 1e-4


def test_adjoint_shares_against_the_oracle(adjoint_state):
    """The gradient at two nodes against the shares built from the oracle's
    MLFMM operator, right-hand side and scattered field at the moved meshes
    (with the I and nu of the GPU solves as the contracted vectors)."""
    from oracle import oracle as O

    a = adjoint_state

    def oracle_at(z):
        p = a["pp"].copy()
        p[:, 2] = z
        o = O.Oracle(p, a["tt"], LAM, 2.25, formulation="ICTF", normal_convention=2)
        o.tree()
        o.near_csr(nthreads=16)
        o.mlfmm_setup(16)
        return o

    for i, nd in enumerate(a["nodes"][:2]):
        zp, zm = a["z0"].copy(), a["z0"].copy()
        zp[nd] += H_FD
        zm[nd] -= H_FD
        op, om = oracle_at(zp), oracle_at(zm)
        dE = op.scattered_field(a["pts"], a["I"])[0].reshape(-1) - \
            om.scattered_field(a["pts"], a["I"])[0].reshape(-1)
        fld = 2 * np.real(np.conj(a["E"]) * a["w"] @ dE)
        rhs = 2 * np.real(a["nu"] @ (op.rhs_plane(a["k"], a["e"]) - om.rhs_plane(a["k"], a["e"])))
        zi = -2 * np.real(a["nu"] @ (op.matvec(a["I"], 16) - om.matvec(a["I"], 16)))
        go = (fld + rhs + zi) / (2 * H_FD)
        print(f"node {nd}: adjoint {a['g'][i]:.6e}, oracle shares {go:.6e}")
        assert abs(a["g"][i] - go) <= 1e-4 * np.max(np.abs(a["g"]))


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
