"""The quadrilateral self term of the restatement
(Integration_KABCD_Taylor_ST_4FEdge, src/lib_sie_quad_singularity.f90:299-380;
oracle taylor_st), pinned against an independent principal-value evaluation.

At an observation point (mu, nu) of an 8-node element, the reference
integrates the five kernels of ff_ABC_II_Edge over the same element in polar
coordinates about the point: 4 angular sectors of ngp = 15 Gauss points, 5
radial points, the 1/rho parts of the strongly singular KA and KB kernels
subtracted (ff_phi_AB_Edge) and added back as log(rho2 |A|) terms.  The KA and
KB integrals are Cauchy principal values: their leading term is odd under
phi -> phi + pi.  With exterior eps_r1 = 1 the KB-E subtraction coefficient
(1 + 1/eps_r2) equals the kernel's own (1/eps_r1 + 1/eps_r2).

The check evaluates the same five integrals with no subtraction and no
analytic term: the rays at phi and phi + pi are added before integrating, so
the odd 1/rho parts cancel pointwise (a bounded integrand, Gauss-Legendre in
rho on the paired part of the rays and in log(rho) beyond it, the angles
split at the corner directions), with the element geometry
from the pinned geometry probe (shape functions, Jacobian, normal, edge
functions and their n.curl; tests/test_quad_oracle.py).  It converges to
1e-6 or better between 24 and 32 points per direction (1e-9 between 32 and
48 at the point nearest a corner).  The reference's rule agrees with it to
3e-8 .. 7e-4 (measured; the largest near a corner), an error set by its fixed
5-point radial rule: it is the same at 15, 30 and 60 angular points.  A wrong sign
of a log term fails by O(1).  (The subtracted 1/rho terms themselves hardly
matter: their radial integral is the same for every ray and their angular
integral vanishes, so a wrong factor there is not visible in the value.)  Then the whole self block (edge_element of an edge against its
own element) is the reference's 15 x 15 outer rule over these points.
"""
import numpy as np
import pytest

from oracle import oracle as O

LAM = 633e-9
EPS1, EPS2 = 1.0, 2.25
K0 = 2 * np.pi / LAM
K1, K2 = K0 * np.sqrt(EPS1), K0 * np.sqrt(EPS2)


@pytest.fixture(scope="module")
def quad():
    nx = ny = 3
    rng = np.random.default_rng(3)
    z = 0.05 * LAM * rng.standard_normal((2 * nx + 1) * (2 * ny + 1))
    return O.QuadOracle(LAM, EPS2, 0.6 * LAM, 0.6 * LAM, nx, ny, z, eps_r1=EPS1)


def _geo(El, k, e):
    g = O.quad_geometry(k, e, El)
    g["r"] = g["N"] @ El
    return g


def _green(k, R):
    ex = np.exp(-1j * k * R)
    return ex / (4 * np.pi * R), -(1 + 1j * k * R) * ex / (4 * np.pi * R ** 2)


def _kernels(El, Ra, Na, mq, nq):
    """The five kernels (KA, KB-E, KB-H, KC-E, KC-H) x 10 source functions at
    the source point (mq, nq), each with the source Jacobian."""
    g = _geo(El, mq, nq)
    R = Ra - g["r"]
    nR = np.linalg.norm(R)
    rn = R / nR
    Ga, DGa = _green(K1, nR)
    Gm, DGm = _green(K2, nR)
    aa = np.cross(g["nq"][None, :], g["fn"])
    f1 = np.cross(aa, rn[None, :]) @ Na * g["JJ"]
    f2 = (Na @ rn) * g["nrot"] * g["JJ"]
    f3 = aa @ Na * g["JJ"]
    return np.stack([f1 * (DGa + DGm), f2 * (DGa / EPS1 + DGm / EPS2), f2 * (DGa + DGm),
                     f3 * (Ga + Gm), f3 * (EPS1 * Ga + EPS2 * Gm)])


def _gl(n, a, b):
    x, w = np.polynomial.legendre.leggauss(n)
    return a + (b - a) * (x + 1) / 2, w * (b - a) / 2


def _gl_log(n, a, b):
    """Gauss-Legendre in log(rho) on [a, b], a > 0: exact for c / rho, the
    leading behaviour of the kernels beyond the paired part of a ray."""
    t, w = _gl(n, 0.0, 1.0)
    L = np.log(b / a)
    r = a * np.exp(t * L)
    return r, w * r * L


def _ray(mu, nu, phi):
    """Distance along direction phi from (mu, nu) to the square's boundary."""
    c, s = np.cos(phi), np.sin(phi)
    t = []
    if c > 1e-15:
        t.append((1 - mu) / c)
    if c < -1e-15:
        t.append((-1 - mu) / c)
    if s > 1e-15:
        t.append((1 - nu) / s)
    if s < -1e-15:
        t.append((-1 - nu) / s)
    return min(t)


def self_point_pv(El, pI, mu, nu, n):
    """Principal values over the element of the five kernels at (mu, nu), by
    paired opposite rays."""
    g0 = _geo(El, mu, nu)
    Ra, Na = g0["r"], g0["fn"][pI - 1]

    def g(rho, phi):
        return _kernels(El, Ra, Na, mu + rho * np.cos(phi), nu + rho * np.sin(phi)) * rho

    corners = [np.arctan2(b - nu, a - mu) % np.pi for a, b in ((1, 1), (-1, 1), (-1, -1), (1, -1))]
    bps = sorted(set([0.0, np.pi] + corners))
    S = np.zeros((5, 10), complex)
    for a, b in zip(bps[:-1], bps[1:]):
        if b - a < 1e-14:
            continue
        for phi, wp in zip(*_gl(n, a, b)):
            ra, rb = _ray(mu, nu, phi), _ray(mu, nu, phi + np.pi)
            m = min(ra, rb)
            I = sum((g(r, phi) + g(r, phi + np.pi)) * w for r, w in zip(*_gl(n, 0.0, m)))
            if ra > m:
                I = I + sum(g(r, phi) * w for r, w in zip(*_gl_log(n, m, ra)))
            if rb > m:
                I = I + sum(g(r, phi + np.pi) * w for r, w in zip(*_gl_log(n, m, rb)))
            S += I * wp
    return S


@pytest.mark.parametrize("elem,pI,mu,nu", [(4, 1, 0.3, -0.2), (4, 7, -0.6, 0.5),
                                           (0, 3, 0.1, 0.05), (8, 10, 0.85, -0.9)])
def test_self_point_equals_principal_value(quad, elem, pI, mu, nu):
    El = quad.element_nodes(elem)
    raw, _ = quad.self_point(El, pI, mu, nu, 15)
    b24 = self_point_pv(El, pI, mu, nu, 24)
    b32 = self_point_pv(El, pI, mu, nu, 32)
    for k in range(5):
        nb = np.linalg.norm(b32[k])
        assert nb > 0
        assert np.linalg.norm(b24[k] - b32[k]) < 1e-6 * nb, (k, "brute force not converged")
        assert np.linalg.norm(raw[k] - b32[k]) < 2e-3 * nb, (k, np.linalg.norm(raw[k] - b32[k]) / nb)


def test_self_block_is_the_outer_rule_over_points(quad):
    """edge_element(s, own element) = sum over the 15 x 15 Gauss points of the
    element (weights times the observation Jacobian) of the self-point rows."""
    ep = quad.edges()
    x, w = np.polynomial.legendre.leggauss(15)
    for s in (1, 9, 20):
        ii, pI = int(ep[s - 1, 0]), int(ep[s - 1, 1])
        El = quad.element_nodes(ii - 1)
        acc = np.zeros((3, 10), complex)
        for j in range(15):
            for i in range(15):
                JJ = O.quad_geometry(x[i], x[j], El)["JJ"]
                acc += quad.self_point(El, pI, x[i], x[j], 15)[1] * (w[i] * w[j] * JJ)
        ref = quad.edge_element(s, ii)
        assert np.linalg.norm(acc - ref) <= 1e-12 * np.linalg.norm(ref), s
