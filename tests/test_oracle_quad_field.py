"""The oracle's quadrilateral scattered field (spq_scattered_field:
Integration_scattered_field_quad, src/lib_sie_quad_calculation_mod.f90:1281-1323,
on the element-local coefficients of lib_sie_quad_solver_normal.f90:312-323),
pinned against the dense fill it shares its kernels with.

With eps_r2 = 1 both media of the fill are the exterior one, so for a test
edge function s and a column tc whose elements share no node with s's
element (the regular 4 x 4 rule, Integration_KABCD_Edge_new), the assembled
PMCHWT entries are twice the exterior field of the unit solution e_tc (or
e_{M+tc}) tested with s's function on its element at the same 4 x 4 points:

  MM[s, tc]       = A  = 2 <N_s, E(e_tc)>      MM[s, M+tc]   = DE = 2 <N_s, E(e_{M+tc})>
  MM[M+s, tc]     = DH = 2 <N_s, H(e_tc)>      MM[M+s, M+tc] = A  = 2 <N_s, H(e_{M+tc})>

which fixes the field's three kernels, their constants and signs and the
element-local coefficients (a wrong sign of the neighbour's share, a
swapped SE/SH or a dropped 1/(i w eps0) fails it).
"""
import numpy as np
import pytest

from oracle import oracle as O
from specklesim_amd import meshes

LAM = 633e-9


@pytest.fixture(scope="module")
def setup():
    Lx, Ly, nx, ny, z = meshes.rough_height_field(5, 5, 1.5, 1.5, 0.05, 0.5, 3, LAM)
    q = O.QuadOracle(LAM, 1.0, Lx, Ly, nx, ny, z)
    MM = q.assemble("PMCHWT", nthreads=8)
    return q, MM


def _test_points(q, s):
    """4 x 4 Gauss points of s's element: positions and w JJ N_s."""
    e = q.edges()
    ii, pI = e[s, 0] - 1, e[s, 1]
    El = q.element_nodes(ii)
    x, w = np.polynomial.legendre.leggauss(4)
    pts, wn = [], []
    for j in range(4):
        for i in range(4):
            g = O.quad_geometry(x[i], x[j], El)
            pts.append(g["N"] @ El)
            wn.append(w[i] * w[j] * g["JJ"] * g["fn"][pI - 1])
    return np.array(pts), np.array(wn)


def test_field_tested_equals_the_fill(setup):
    q, MM = setup
    M = q.M
    e = q.edges()
    vert = q.elements()
    # s on the corner element 0, columns on elements with no node in common
    rows = [s for s in range(M) if e[s, 0] == 1][:4]
    near = set(vert[0])
    far_cols = [tc for tc in range(M)
                if not (set(vert[e[tc, 0] - 1]) & near)
                and (e[tc, 2] == 0 or not (set(vert[e[tc, 2] - 1]) & near))]
    rng = np.random.default_rng(1)
    cols = rng.choice(far_cols, 6, replace=False)
    # at least one column shared by two elements with the sign of the neighbour
    assert any(e[tc, 2] > 0 for tc in cols)
    worst = 0.0
    for s in rows:
        pts, wn = _test_points(q, s)
        for tc in cols:
            for cur in (0, 1):
                x = np.zeros(2 * M, complex)
                x[cur * M + tc] = 1.0
                E, H = q.scattered_field(pts, x, ngp=4)
                tE = 2 * np.sum(wn * E)
                tH = 2 * np.sum(wn * H)
                for got, ref in ((tE, MM[s, cur * M + tc]), (tH, MM[M + s, cur * M + tc])):
                    worst = max(worst, abs(got - ref) / abs(ref))
    print("worst relative difference", worst)
    assert worst < 1e-12


def test_plane_wave_field_decays_as_a_far_field(setup):
    """A solved problem's field at r and 2r far from the patch falls as 1/r
    in amplitude (|E| r constant to the far-field order)."""
    q, MM = setup
    b = q.rhs_plane([0, 0, -1.0], [1.0, 0, 0], "PMCHWT")
    x = np.linalg.solve(MM, b)
    d = np.array([0.3, 0.2, 0.93])
    d /= np.linalg.norm(d)
    r1, r2 = 2e3 * LAM, 4e3 * LAM
    E1, _ = q.scattered_field(d[None, :] * r1, x)
    E2, _ = q.scattered_field(d[None, :] * r2, x)
    a1, a2 = np.linalg.norm(E1) * r1, np.linalg.norm(E2) * r2
    assert abs(a1 - a2) / a2 < 1e-3
