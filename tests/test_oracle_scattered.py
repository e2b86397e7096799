"""Pins the restatement of Integration_scattered_field_tri
(src/lib_sie_tri_calculation_mod.f90:2782-2825) with closed forms: an RWG
pair of size d = 1e-7 lambda radiates, at one wavelength, as a point dipole.
The 3-point rule integrates f exactly, so the only error is the next
multipole order, relative size O(k d) ~ 1e-6.

Electric current J: p = J * int f dS = J * l (c- - c+),
  E = -i w mu G p  (vector-potential part only: the reference multiplies the
                    charge term of E by 0),
  H = grad G x p.
Magnetic current M: p_m likewise,
  E = p_m x grad G,
  H = -i w eps G p_m + (p_m . grad) grad G / (i w mu).
"""
import numpy as np

from oracle import oracle as O

LAM = 633e-9


def _tiny_pair(scale):
    pp = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.3, 0.9, 0.0], [0.8, -0.7, 0.2]]) * scale
    tt = np.array([[1, 2, 3], [2, 1, 4]], np.int32)
    return pp, tt


def _green(k, R):
    e = np.exp(-1j * k * R)
    G = e / (4 * np.pi * R)
    G1 = -(1 + 1j * k * R) * e / (4 * np.pi * R ** 2)
    G2 = e * (2 + 2j * k * R - (k * R) ** 2) / (4 * np.pi * R ** 3)
    return G, G1, G2


def test_scattered_field_is_dipole_field():
    pp, tt = _tiny_pair(1e-7 * LAM)
    o = O.Oracle(pp, tt, LAM, 2.25, formulation="PMCHWT", normal_convention=1)
    assert o.N == 1
    # records: p1, p2, pc (9), area, length, df, mid (3), nq (3), sign
    f1, f2 = o.fn(0, 1), o.fn(0, 2)
    c = [f[:9].reshape(3, 3) for f in (f1, f2)]
    cen = [ci.mean(axis=0) for ci in c]
    s = [f[18] for f in (f1, f2)]
    l = f1[10]
    # int f dS = sum_t sign_t (l/2)(centroid_t - v_t)
    pint = sum(s[t] * 0.5 * l * (cen[t] - c[t][2]) for t in range(2))
    # the reference's single-precision constants (lib_sie_constants.f90)
    k = 2 * np.pi / LAM
    omega = 2 * np.pi * float(np.float32(2.99792458e8)) / LAM
    mu = float(np.float32(1.2566370614e-6))
    eps = float(np.float32(8.854187817e-12))
    rng = np.random.default_rng(3)
    pts = rng.standard_normal((5, 3))
    pts = pts / np.linalg.norm(pts, axis=1)[:, None] * LAM
    for J, M in ((1.0, 0.0), (0.0, 1.0)):
        E, H = o.scattered_field(pts, np.array([J, M], complex))
        for i, ra in enumerate(pts):
            R = np.linalg.norm(ra)
            rh = ra / R
            G, G1, G2 = _green(k, R)
            gG = rh * G1
            if J:
                p = J * pint
                Ee = -1j * omega * mu * G * p
                He = np.cross(gG, p)
            else:
                p = M * pint
                Ee = np.cross(p, gG)
                hess_p = p * G1 / R + rh * np.dot(rh, p) * (G2 - G1 / R)
                He = -1j * omega * eps * G * p + hess_p / (1j * omega * mu)
            assert np.linalg.norm(E[i] - Ee) < 1e-5 * np.linalg.norm(Ee), (J, M, i)
            assert np.linalg.norm(H[i] - He) < 1e-5 * np.linalg.norm(He), (J, M, i)


def test_scattered_field_linear_in_currents():
    pp, tt = _tiny_pair(0.3 * LAM)
    o = O.Oracle(pp, tt, LAM, 2.25)
    pts = np.array([[0.1, 0.2, 1.0], [2.0, -1.0, 0.5]]) * LAM
    x1, x2 = np.array([1.0 + 2j, 0.5]), np.array([-0.3j, 1.0 - 1j])
    E1, H1 = o.scattered_field(pts, x1)
    E2, H2 = o.scattered_field(pts, x2)
    E3, H3 = o.scattered_field(pts, 2 * x1 - 1j * x2)
    assert np.allclose(E3, 2 * E1 - 1j * E2, rtol=1e-13, atol=0)
    assert np.allclose(H3, 2 * H1 - 1j * H2, rtol=1e-13, atol=0)
