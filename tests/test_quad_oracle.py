"""Pins the quadrilateral (8-node, 10 Huber edge functions) restatement in
oracle/spk_quad.c and checks the library's host topology against it.

What pins what:
* element geometry: partition of unity and nodal interpolation of the
  shape functions, their derivatives against finite differences, the
  Jacobian against finite-difference tangents, and n.curl of each edge
  function (n_rot_N, src/lib_sie_quad_calculation_mod.f90:436-450) against a
  finite-difference curl of its covariant components;
* regular and corner-adjacent pairs: the rule is a plain Gauss product, so
  the reference's 4- and 6-point results equal the probe at that order and
  converge with the order;
* edge-adjacent pairs (GL_Int_*_EA_8Node): brute-force Gauss products of
  rising order approach the reference's singular rule monotonically;
* self pairs (Integration_KABCD_Taylor_ST_4FEdge): no independent value
  (source and observation points coincide) -- parity unpinned beyond the
  GPU agreeing with this restatement;
* numbering (Sorting_edge): every (element, local function) appears once,
  M = 2 n_el + 2 (number of element sides).
"""
import numpy as np
import pytest

from oracle import oracle as O
from specklesim_amd import meshes

LAM = 633e-9
NODES_UV = [(-1, -1), (0, -1), (1, -1), (1, 0), (1, 1), (0, 1), (-1, 1), (-1, 0)]


def curved_element():
    pts = []
    for (u, v) in NODES_UV:
        x, y = 0.8 * u + 0.1 * v, 0.6 * v - 0.05 * u * u
        pts.append((x, y, 0.3 * x * x - 0.2 * y * y + 0.1 * x * y))
    return np.array(pts) * LAM


def test_shape_functions():
    El = curved_element()
    for i, (u, v) in enumerate(NODES_UV):
        g = O.quad_geometry(u, v, El)
        assert np.allclose(g["N"], np.eye(8)[i], atol=1e-15)
    rng = np.random.default_rng(1)
    h = 1e-6
    for _ in range(10):
        k, e = rng.uniform(-1, 1, 2)
        g = O.quad_geometry(k, e, El)
        assert abs(g["N"].sum() - 1) < 1e-14
        gk = (O.quad_geometry(k + h, e, El)["N"] - O.quad_geometry(k - h, e, El)["N"]) / (2 * h)
        ge = (O.quad_geometry(k, e + h, El)["N"] - O.quad_geometry(k, e - h, El)["N"]) / (2 * h)
        assert np.allclose(g["dk"], gk, atol=1e-8) and np.allclose(g["de"], ge, atol=1e-8)


def _r(k, e, El):
    return O.quad_geometry(k, e, El)["N"] @ El


def test_jacobian_normal_and_curl_of_edge_functions():
    El = curved_element()
    rng = np.random.default_rng(2)
    h = 1e-5
    for _ in range(8):
        k, e = rng.uniform(-0.9, 0.9, 2)
        g = O.quad_geometry(k, e, El)
        ru = (_r(k + h, e, El) - _r(k - h, e, El)) / (2 * h)
        rv = (_r(k, e + h, El) - _r(k, e - h, El)) / (2 * h)
        n = np.cross(ru, rv)
        J = np.linalg.norm(n)
        assert abs(g["JJ"] / J - 1) < 1e-8
        assert np.allclose(g["nq"], n / J, atol=1e-8)

        def cov(kk, ee):
            gg = O.quad_geometry(kk, ee, El)
            u_ = (_r(kk + h, ee, El) - _r(kk - h, ee, El)) / (2 * h)
            v_ = (_r(kk, ee + h, El) - _r(kk, ee - h, El)) / (2 * h)
            return gg["fn"] @ u_, gg["fn"] @ v_  # covariant components N.r_u, N.r_v

        d = 1e-4
        Nu_pe, _ = cov(k, e + d)
        Nu_me, _ = cov(k, e - d)
        _, Nv_pk = cov(k + d, e)
        _, Nv_mk = cov(k - d, e)
        curl = ((Nv_pk - Nv_mk) - (Nu_pe - Nu_me)) / (2 * d) / J
        assert np.allclose(g["nrot"], curl, rtol=1e-5, atol=1e-6 * np.abs(g["nrot"]).max())


@pytest.fixture(scope="module")
def q3():
    Lx, Ly, nx, ny, z = meshes.rough_height_field(3, 3, 1.2, 1.2, 0.05, 0.5, 2, LAM)
    return O.QuadOracle(LAM, 2.25, Lx, Ly, nx, ny, z), (Lx, Ly, nx, ny, z)


def _edge_fn_on(q, elem):
    e = q.edges()
    return [i + 1 for i in range(q.M) if e[i, 0] == elem]


def test_numbering(q3):
    q, (Lx, Ly, nx, ny, z) = q3
    e = q.edges()
    sides = nx * (ny + 1) + ny * (nx + 1)
    assert q.M == 2 * q.n_el + 2 * sides
    seen = {}
    for (ii, pa, jj, pq, f) in e:
        for key in [(ii, pa)] + ([(jj, pq)] if f != 0 else []):
            assert key not in seen
            seen[key] = 1
        assert f in (-1, 0, 1)
        if pa in (5, 10):
            assert f == 0  # the interior functions are not shared
    assert len(seen) == 10 * q.n_el


def test_regular_and_corner_rules(q3):
    q, _ = q3
    s = _edge_fn_on(q, 5)[0]          # centre element of the 3 x 3 grid
    # element 1 touches element 5 at a corner, a far pair needs a 4 x 4 grid:
    # the corner rule is the 6-point product rule of the probe
    ref = q.edge_element(s, 1)
    assert np.array_equal(ref, q.regular_pair(s, 1, 6))
    hi = q.regular_pair(s, 1, 20)
    assert np.linalg.norm(q.regular_pair(s, 1, 16) - hi) < 1e-5 * np.linalg.norm(hi)
    assert np.linalg.norm(ref - hi) < 1e-3 * np.linalg.norm(hi)


def test_far_pair_rule():
    Lx, Ly, nx, ny, z = meshes.rough_height_field(4, 1, 1.6, 0.4, 0.05, 0.5, 3, LAM)
    q = O.QuadOracle(LAM, 2.25, Lx, Ly, nx, ny, z)
    s = _edge_fn_on(q, 1)[0]
    ref = q.edge_element(s, 4)       # no shared node
    assert np.array_equal(ref, q.regular_pair(s, 4, 4))
    hi = q.regular_pair(s, 4, 16)
    assert np.linalg.norm(ref - hi) < 1e-3 * np.linalg.norm(hi)


def test_edge_adjacent_rule_against_brute_force(q3):
    q, _ = q3
    s = _edge_fn_on(q, 5)[2]
    ref = q.edge_element(s, 2)       # element 2 shares a side with element 5
    d = []
    for ng in (16, 32, 48):
        r = q.regular_pair(s, 2, ng)
        d.append([np.linalg.norm(r[k] - ref[k]) / np.linalg.norm(ref[k]) for k in range(3)])
    d = np.array(d)
    assert np.all(np.diff(d, axis=0) < 0)  # brute force closes in on the singular rule
    assert d[-1, 0] < 5e-3 and np.all(d[-1, 1:] < 0.1)


def test_library_host_topology_matches(q3):
    from specklesim_amd import quad

    q, (Lx, Ly, nx, ny, z) = q3
    pp, v, e = quad.topology_probe(Lx, Ly, nx, ny, z)
    assert np.array_equal(pp, q.nodes())
    assert np.array_equal(v, q.elements())
    assert np.array_equal(e, q.edges())
    for (a, b) in ((5, 2), (1, 6), (2, 2)):
        Lx, Ly, nx, ny, z = meshes.rough_height_field(a, b, 0.4 * a, 0.4 * b, 0.1, 0.5, a + b, LAM)
        qo = O.QuadOracle(LAM, 2.25, Lx, Ly, nx, ny, z)
        pp, v, e = quad.topology_probe(Lx, Ly, nx, ny, z)
        assert np.array_equal(pp, qo.nodes()) and np.array_equal(v, qo.elements())
        assert np.array_equal(e, qo.edges())


def test_height_file_round_trip(tmp_path):
    Lx, Ly, nx, ny, z = meshes.rough_height_field(3, 2, 1.0, 0.7, 0.05, 0.5, 5, LAM)
    f = tmp_path / "surface.txt"
    meshes.write_quad_height_file(str(f), Lx, Ly, nx, ny, z)
    Lx2, Ly2, nx2, ny2, z2 = meshes.read_quad_height_file(str(f))
    assert (Lx2, Ly2, nx2, ny2) == (Lx, Ly, nx, ny) and np.array_equal(z2, z)


def _brute_rhs(q, k, E0, ng=24):
    """int over element edge_p(1) of N_pI . E_inc, N_pI . H_inc by a ng x ng
    Gauss product, from the pinned geometry probe (lossless exterior)."""
    x, w = np.polynomial.legendre.leggauss(ng)
    pp, v, e = q.nodes(), q.elements(), q.edges()
    k1 = 2 * np.pi / LAM
    mu0c0 = float(np.float32(1.2566370614e-6)) * float(np.float32(2.99792458e8))
    out = np.zeros((q.M, 2), complex)
    for s in range(q.M):
        El = pp[v[e[s, 0] - 1] - 1]
        pI = e[s, 1] - 1
        for j in range(ng):
            for i in range(ng):
                g = O.quad_geometry(x[i], x[j], El)
                R = g["N"] @ El
                E = E0 * np.exp(-1j * k1 * (k @ R))
                H = np.cross(k, E) / mu0c0
                t = w[i] * w[j] * g["JJ"]
                out[s, 0] += (g["fn"][pI] @ E) * t
                out[s, 1] += (g["fn"][pI] @ H) * t
    return out


def test_plane_wave_rhs_against_brute_force():
    """get_vector_b_quad (src/lib_sie_quad_calculation_mod.f90:91-138): each
    edge function tested on its own element only, with the 6-point rule; a
    24-point product of the same integrand agrees to the 6-point rule's error."""
    Lx, Ly, nx, ny, z = meshes.rough_height_field(2, 2, 0.8, 0.8, 0.05, 0.5, 4, LAM)
    q = O.QuadOracle(LAM, 2.25, Lx, Ly, nx, ny, z)
    th = np.deg2rad(30.0)
    k = np.array([np.sin(th), 0.0, -np.cos(th)])
    E0 = np.array([np.cos(th), 0.0, np.sin(th)])
    b = q.rhs_plane(k, E0, "PMCHWT")
    ref = _brute_rhs(q, k, E0)
    M = q.M
    assert np.linalg.norm(b[:M] - ref[:, 0]) < 1e-5 * np.linalg.norm(ref[:, 0])
    assert np.linalg.norm(b[M:] - ref[:, 1]) < 1e-5 * np.linalg.norm(ref[:, 1])


def test_plane_wave_rhs_height_shift_phase():
    """Raising the surface by h multiplies the RHS by exp(-i conj(k1) k_z h):
    the phase of dot_product(k1*k_in, R_a) (:1231), whose complex first
    argument Fortran conjugates (visible with a lossy exterior medium)."""
    Lx, Ly, nx, ny, z = meshes.rough_height_field(2, 3, 0.8, 1.1, 0.05, 0.5, 8, LAM)
    eps1 = 1.7 - 0.3j
    h = 0.41 * LAM
    k = np.array([0.6, 0.0, -0.8])
    E0 = np.array([0.8, 0.0, 0.6])
    q0 = O.QuadOracle(LAM, 2.25, Lx, Ly, nx, ny, z, eps_r1=eps1)
    q1 = O.QuadOracle(LAM, 2.25, Lx, Ly, nx, ny, z + h, eps_r1=eps1)
    k1 = 2 * np.pi / LAM * np.sqrt(eps1)
    f = np.exp(-1j * np.conj(k1) * k[2] * h)
    assert abs(f - np.exp(-1j * k1 * k[2] * h)) > 0.1 * abs(f)  # the test tells them apart
    for form in ("PMCHWT", "ICTF"):
        b0, b1 = q0.rhs_plane(k, E0, form), q1.rhs_plane(k, E0, form)
        assert np.linalg.norm(b1 - b0 * f) < 1e-12 * np.linalg.norm(b1)
