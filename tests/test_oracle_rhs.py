"""Pins the plane-wave right-hand side of the oracle (get_vector_b_tri +
incident_field_tri_general, src/lib_sie_tri_calculation_mod.f90:391-460,
939-1002) without restating it:

* constant field (normal incidence on a flat patch at z = 0): the rule is
  exact for the linear RWG function (to the single-precision rounding of
  its constants, ~1e-8), so each of the four integrals is the
  closed form  v . int f_m dS  with  int f_m dS = sum_t sign_t (l/2)(c_t - v_t),
  v = E0, H0 = (k x E0)/(mu1 c0), n x E0, n x H0 -- seen through PMCHWT
  (E, H) and CNF (n x H, -n x E);
* translation: moving the mesh by d multiplies every entry by
  exp(-i conj(k1) k.d) -- the phase of dot_product(k1*k_in, R_a) (:977),
  whose complex first argument Fortran conjugates; a lossy exterior medium
  makes the conjugation visible (|factor| = exp(-Im(k1) k.d) > 1 for k.d > 0).
"""
import numpy as np

from oracle import oracle as O
from specklesim_amd import meshes

LAM = 633e-9
MU0 = float(np.float32(1.2566370614e-6))
C0 = float(np.float32(2.99792458e8))


def _flat_patch():
    pp, tt = meshes.rough_surface(4, 3, 1.3, 1.0, 0.05, 0.5, 7, LAM)
    pp = pp.copy()
    pp[:, 2] = 0.0
    return pp, tt


def _int_f(o, m):
    s = np.zeros(3)
    for ot in (1, 2):
        r = o.fn(m, ot)
        c = r[:9].reshape(3, 3)
        s += r[18] * 0.5 * r[10] * (c.mean(axis=0) - c[2])
    return s


def test_constant_field_closed_form():
    pp, tt = _flat_patch()
    k = np.array([0.0, 0.0, -1.0])
    E0 = np.array([0.6, -0.8, 0.0])
    H0 = np.cross(k, E0) / (MU0 * C0)
    for form in ("PMCHWT", "CNF"):
        o = O.Oracle(pp, tt, LAM, 2.25, formulation=form)
        b = o.rhs_plane(k, E0)
        N = o.N
        for m in range(N):
            fI = _int_f(o, m)
            nq = o.fn(m, 1)[15:18]
            assert np.allclose(o.fn(m, 2)[15:18], nq)  # flat: one normal
            if form == "PMCHWT":
                ref = (E0 @ fI, H0 @ fI)
            else:
                ref = (np.cross(nq, H0) @ fI, -(np.cross(nq, E0) @ fI))
            # the reference's 3-point rule constants are single-precision
            # literals, so the rule integrates linear functions to ~1e-8
            sE = np.linalg.norm(E0) * np.linalg.norm(fI)
            sH = sE / (MU0 * C0)
            scale = (sE, sH) if form == "PMCHWT" else (sH, sE)
            assert abs(b[m] - ref[0]) <= 1e-7 * scale[0], (form, m)
            assert abs(b[m + N] - ref[1]) <= 1e-7 * scale[1], (form, m)


def test_translation_phase_uses_conjugated_k1():
    pp, tt = meshes.rough_surface(3, 3, 1.0, 1.0, 0.05, 0.5, 2, LAM)
    eps1 = 1.7 - 0.3j
    k = np.array([0.6, 0.0, -0.8])
    E0 = np.array([0.8, 0.0, 0.6])
    d = np.array([0.37, -0.21, 0.05]) * LAM
    o0 = O.Oracle(pp, tt, LAM, 2.25, eps_r1=eps1)
    o1 = O.Oracle(pp + d, tt, LAM, 2.25, eps_r1=eps1)
    b0, b1 = o0.rhs_plane(k, E0), o1.rhs_plane(k, E0)
    k1 = 2 * np.pi / LAM * np.sqrt(eps1)
    factor = np.exp(-1j * np.conj(k1) * (k @ d))
    assert abs(factor) > 1.1  # conjugation grows the wave along k
    assert np.linalg.norm(b1 - b0 * factor) < 1e-12 * np.linalg.norm(b1)
