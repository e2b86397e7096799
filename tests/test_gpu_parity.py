"""GPU parity: libspeckle_gpu.so (HIP, gfx950) against the CPU restatement on
the same inputs, through the C-ABI.

Integer outputs (RWG pairs, octree leaves, CSR structure, I4 lists) are
compared bit-exactly.  Floating point: near-field blocks per component to
1e-10 (norm-wise) -- the reference is FP64 and the GPU differs only by
evaluation order and FMA contraction; the MLFMM matvec to 1e-10 relative L2
on geometries whose leaf boxes are >= 0.6 wavelength (smaller boxes make
the K=19 translation sum ill-conditioned in the reference algorithm itself:
rounding noise ~1e-16 * |TL| ~ 1e-5, so parity there is stated as the
conditioning bound); GMRES solutions to the conditioning of the solve.
"""
import numpy as np
import pytest

from oracle import oracle as O
from specklesim_amd import DeviceBuffer, SieTriMlfmm, meshes, plane_wave

pytestmark = pytest.mark.gpu
LAM = 633e-9
FORMS = ["PMCHWT", "MCTF", "ICTF", "MCTF2", "PMCHWT-S", "JMCFIE", "CNF", "MNMF"]


def rel(a, b):
    nb = np.linalg.norm(b)
    if nb == 0.0:  # a block component that is identically zero (CNF's Z_hh)
        return np.linalg.norm(a)
    return np.linalg.norm(a - b) / nb


@pytest.fixture(scope="module")
def patch5():
    """5 lambda x 5 lambda rough patch, 50x50 cells: N = 7400, l_max = 3,
    0.625 lambda leaf boxes (the leaf size of the 10 lambda workload)."""
    pp, tt = meshes.rough_surface(50, 50, 5.0, 5.0, 0.1, 0.5, 3, LAM)
    o = O.Oracle(pp, tt, LAM, 2.25, formulation="PMCHWT", normal_convention=2)
    o.tree()
    ia, ja, v = o.near_csr(nthreads=16)
    o.mlfmm_setup(16)
    g = SieTriMlfmm(LAM, 2.25, formulation="PMCHWT", normal_convention=2).precalculation_tri(pp, tt)
    return pp, tt, o, g, (ia, ja, v)


def test_integer_structures_bit_exact(patch5):
    pp, tt, o, g, (ia, ja, v) = patch5
    c1, e1 = g.rwg()
    c2, e2 = o.pairs()
    assert np.array_equal(c1, c2) and np.array_equal(e1, e2)
    tg, to = g.tree(), o.tree()
    assert (tg["l_min"], tg["l_max"]) == (to["l_min"], to["l_max"]) == (2, 3)
    assert np.array_equal(tg["leaf"], to["leaf"])
    ia2, ja2, _ = g.near_csr(values=False)
    assert np.array_equal(ia2, ia) and np.array_equal(ja2, ja)
    for level in (2, 3):
        for box in np.unique(to["leaf"] >> (3 * (3 - level)))[::7]:
            ref = [b for b in o.i4(level, int(box)) if b in set((to["leaf"] >> (3 * (3 - level))).tolist())]
            assert np.array_equal(g.i4(level, int(box)), np.array(sorted(ref)))


def test_near_field_blocks(patch5):
    pp, tt, o, g, (ia, ja, v) = patch5
    _, _, vg = g.near_csr()
    for k in range(4):
        assert rel(vg[:, k], v[:, k]) < 1e-10, k
    # every block row to 1e-9 of its own scale
    scale = np.abs(v).max(axis=1)
    assert np.max(np.abs(vg - v).max(axis=1) / scale) < 1e-9


def test_matvec(patch5):
    pp, tt, o, g, _ = patch5
    rng = np.random.default_rng(7)
    x = rng.standard_normal(2 * o.N) + 1j * rng.standard_normal(2 * o.N)
    yo = o.matvec(x, 16)
    yg = g.calculate_vector_b_tri(x)
    assert rel(yg[:o.N], yo[:o.N]) < 1e-10
    assert rel(yg[o.N:], yo[o.N:]) < 1e-10


def test_patterns(patch5):
    pp, tt, o, g, _ = patch5
    P = g.patterns(0, 20)
    for m in range(20):
        assert rel(P[m], o.pattern(m)) < 1e-12


def test_rhs(patch5):
    pp, tt, o, g, _ = patch5
    k, e = plane_wave(0.3, 0.2, 2, "TM")
    b = g.get_vector_b_tri(k, e)
    assert rel(b, o.rhs_plane(k, e)) < 1e-13
    d_b = DeviceBuffer(2 * g.N * 16)
    g.get_vector_b_tri_device(k, e, d_b.ptr)
    assert np.array_equal(d_b.download(), b)


def test_matvec_linear_and_deterministic(patch5):
    _, _, o, g, _ = patch5
    rng = np.random.default_rng(8)
    x1 = rng.standard_normal(2 * o.N) + 1j * rng.standard_normal(2 * o.N)
    x2 = rng.standard_normal(2 * o.N) + 1j * rng.standard_normal(2 * o.N)
    a, b = 0.3 - 1.1j, 2.0 + 0.5j
    y = g.calculate_vector_b_tri(a * x1 + b * x2)
    assert rel(y, a * g.calculate_vector_b_tri(x1) + b * g.calculate_vector_b_tri(x2)) < 1e-12
    assert np.array_equal(g.calculate_vector_b_tri(x1), g.calculate_vector_b_tri(x1))
    yn, yf = g.matvec_parts(x1)
    assert rel(yn + yf, g.calculate_vector_b_tri(x1)) < 1e-14


def test_scattered_field_and_brdf(patch5):
    """Integration_scattered_field_tri on the GPU against the restatement:
    near-zone and far-zone points, host and device x; the BRDF curve."""
    from specklesim_amd import farfield

    pp, tt, o, g, _ = patch5
    rng = np.random.default_rng(11)
    x = rng.standard_normal(2 * o.N) + 1j * rng.standard_normal(2 * o.N)
    pts = np.concatenate([rng.standard_normal((6, 3)) * 3 * LAM,
                          farfield.sphere_points((-1.2, 1.2), 9, 0.3, 1e3 * LAM)[1]])
    Eg, Hg = g.scattered_field(pts, x)
    Eo, Ho = o.scattered_field(pts, x)
    for i in range(pts.shape[0]):
        assert np.linalg.norm(Eg[i] - Eo[i]) < 1e-12 * np.linalg.norm(Eo[i]), i
        assert np.linalg.norm(Hg[i] - Ho[i]) < 1e-12 * np.linalg.norm(Ho[i]), i
    d_x = DeviceBuffer(x.nbytes).upload(x)
    Ed, Hd = g.scattered_field_device(pts, d_x.ptr)
    assert np.array_equal(Ed, Eg) and np.array_equal(Hd, Hg)
    th, bg = farfield.brdf(g, x, 0.0, (-1.2, 1.2), 9, 0.3, 1e3 * LAM)
    _, po = farfield.sphere_points((-1.2, 1.2), 9, 0.3, 1e3 * LAM)
    bo = farfield.intensity(o.scattered_field(po, x)[0], po) / np.cos(th)
    assert rel(bg, bo) < 1e-12


@pytest.mark.parametrize("form", FORMS)
def test_all_formulations_near_blocks(form):
    pp, tt = meshes.rough_surface(12, 12, 1.2, 1.2, 0.1, 0.5, 9, LAM)
    eps2 = 18.392 - 0.4164j  # lossy interior (silicon at 500 nm, main.f90:133)
    o = O.Oracle(pp, tt, LAM, eps2, formulation=form, normal_convention=2, s_opt=30)
    g = SieTriMlfmm(LAM, eps2, formulation=form, normal_convention=2,
                    tree_s_opt=30).precalculation_tri(pp, tt)
    rng = np.random.default_rng(5)
    m = rng.integers(0, o.N, 300)
    n = rng.integers(0, o.N, 300)
    # include self blocks and edge neighbours
    m = np.concatenate([m, np.arange(40)])
    n = np.concatenate([n, np.arange(40)])
    zg = g.impedance_edges(m, n)
    zo = np.array([o.near_block(int(a), int(b)) for a, b in zip(m, n)])
    for i in range(2):
        for j in range(2):
            assert rel(zg[:, i, j], zo[:, i, j]) < 1e-10, (form, i, j)
    # the formulation's right-hand side (oblique TE incidence)
    k, e = plane_wave(0.4, 0.9, 2, "TE")
    assert rel(g.get_vector_b_tri(k, e), o.rhs_plane(k, e)) < 1e-13, form
    # the fill kernels (classified task lists) equal the pair probe
    ia, ja, vg = g.near_csr()
    rows = np.repeat(np.arange(o.N), np.diff(ia))
    pick = np.arange(0, ja.size, 97)
    zp = g.impedance_edges(rows[pick], ja[pick])
    assert rel(vg[pick].reshape(-1, 2, 2), zp) < 1e-13


def _patch24(form, s_opt):
    pp, tt = meshes.rough_surface(24, 24, 2.4, 2.4, 0.1, 0.5, 6, LAM)
    o = O.Oracle(pp, tt, LAM, 2.25, formulation=form, normal_convention=2, s_opt=s_opt)
    o.tree()
    o.near_csr(nthreads=16)
    o.mlfmm_setup(16)
    g = SieTriMlfmm(LAM, 2.25, formulation=form, normal_convention=2,
                    tree_s_opt=s_opt).precalculation_tri(pp, tt)
    k, e = plane_wave(0.0, 0.0, 2, "TM")
    return o, g, g.get_vector_b_tri(k, e), o.rhs_plane(k, e)


@pytest.mark.parametrize("form", FORMS)
def test_all_formulations_matvec(form):
    """The whole matvec of every formulation: its far part carries the
    formulation's scaling of the disaggregated sums
    (src/lib_sie_tri_mlfmm.f90:478-495)."""
    o, g, _, _ = _patch24(form, 250)
    rng = np.random.default_rng(11)
    x = rng.standard_normal(2 * o.N) + 1j * rng.standard_normal(2 * o.N)
    yo = o.matvec(x, 16)
    yg = g.calculate_vector_b_tri(x)
    N = o.N
    assert rel(yg[:N], yo[:N]) < 1e-10, form
    assert rel(yg[N:], yo[N:]) < 1e-10, form


def test_gmres_converged_solution():
    """ICTF on the 2.4 lambda patch (N = 1680, one far-field level, 0.58
    lambda leaves): GMRES converges to 1e-8 in ~360 iterations on both."""
    o, g, b, bo = _patch24("ICTF", 250)
    xg, ig = g.gmres_run(b, tolerance=1e-8, max_iterations=400)
    xo, io = o.gmres(bo, tol=1e-8, maxit=400, nthreads=16)
    assert ig["info"] == io["info"] == 0
    assert abs(ig["iters"] - io["iters"]) <= 2
    n = min(len(ig["hist"]), len(io["hist"]))
    assert np.max(np.abs(np.log(ig["hist"][:n] / io["hist"][:n]))) < 1e-3
    assert rel(xg, xo) < 1e-6
    # the returned backward error is the true residual of the returned x
    r = b - g.calculate_vector_b_tri(xg)
    assert abs(np.linalg.norm(r) / np.linalg.norm(b) - ig["be"]) < 1e-10
    # and the GPU solution solves the restatement's operator to the tolerance
    ro = bo - o.matvec(xg, 16)
    assert np.linalg.norm(ro) / np.linalg.norm(b) < 1.01e-8


def test_gmres_left_jacobi():
    """The reference's Jacobi left preconditioner (precondition_left) on
    PMCHWT, which stagnates near 1.5e-3 without it on this open patch:
    GPU and restatement converge together; the second backward error is the
    unpreconditioned residual."""
    o, g, b, bo = _patch24("PMCHWT", 250)
    xg, ig = g.gmres_run(b, tolerance=1e-8, max_iterations=400, precondition_left=True)
    xo, io = o.gmres_left_jacobi(bo, tol=1e-8, maxit=400, nthreads=16)
    assert ig["info"] == io["info"] == 0
    assert abs(ig["iters"] - io["iters"]) <= 2
    n = min(len(ig["hist"]), len(io["hist"]))
    assert np.max(np.abs(np.log(ig["hist"][:n] / io["hist"][:n]))) < 1e-3
    assert rel(xg, xo) < 1e-5
    r = b - g.calculate_vector_b_tri(xg)
    assert abs(np.linalg.norm(r) / np.linalg.norm(b) - ig["tbe"]) < 1e-10
    assert abs(ig["tbe"] / io["tbe"] - 1) < 1e-2


def test_gmres_two_level_history():
    """PMCHWT on the same patch with s_opt = 60 (two far-field levels, 0.29
    lambda leaves; PMCHWT on an open patch stagnates near 1.5e-3): 100
    Arnoldi steps; the K = 19 sum's ~1e-8 rounding noise bounds agreement."""
    o, g, b, bo = _patch24("PMCHWT", 60)
    xg, ig = g.gmres_run(b, tolerance=1e-12, max_iterations=100)
    xo, io = o.gmres(bo, tol=1e-12, maxit=100, nthreads=16)
    assert ig["info"] == io["info"] == -4 and ig["iters"] == io["iters"] == 100
    assert np.max(np.abs(np.log(ig["hist"] / io["hist"]))) < 1e-4
    assert rel(xg, xo) < 1e-4


@pytest.mark.parametrize("orth", ["MGS", "IMGS", "CGS", "ICGS"])
def test_gmres_orthogonalization_schemes(patch5, orth):
    """Each Gram-Schmidt variant of the GMRES host against the restatement's
    same variant: 60 Arnoldi steps on the 7400-edge PMCHWT patch."""
    pp, tt, o, g, _ = patch5
    k, e = plane_wave(0.0, 0.0, 2, "TM")
    b = g.get_vector_b_tri(k, e)
    x, info = g.gmres_run(b, tolerance=1e-12, max_iterations=60, orthogonalization=orth)
    code = {"MGS": 0, "IMGS": 1, "CGS": 2, "ICGS": 3}[orth]
    xo, io = o.gmres(o.rhs_plane(k, e), tol=1e-12, maxit=60, orth=code, nthreads=16)
    assert info["info"] == io["info"] == -4 and info["iters"] == io["iters"] == 60
    assert np.max(np.abs(np.log(info["hist"] / io["hist"]))) < 1e-6
    assert abs(info["be"] / io["be"] - 1) < 1e-6
    assert rel(x, xo) < 1e-6


@pytest.mark.parametrize("kind", ["sweep", "steps"])
def test_mgs_pass_forms_agree(patch5, kind, monkeypatch):
    """The three forms of one MGS pass -- one-reduce (default: h = (I + L)^-1
    V^H w), the cooperative column sweep and one launch pair per column --
    give the same IMGS history and solution to rounding."""
    pp, tt, o, g, _ = patch5
    k, e = plane_wave(0.0, 0.0, 2, "TM")
    b = g.get_vector_b_tri(k, e)
    monkeypatch.delenv("SPK_MGS", raising=False)
    x0, i0 = g.gmres_run(b, tolerance=1e-12, max_iterations=40, orthogonalization="IMGS")
    monkeypatch.setenv("SPK_MGS", kind)
    x1, i1 = g.gmres_run(b, tolerance=1e-12, max_iterations=40, orthogonalization="IMGS")
    assert i0["iters"] == i1["iters"] == 40
    assert np.max(np.abs(np.log(i0["hist"] / i1["hist"]))) < 1e-8
    assert rel(x1, x0) < 1e-8


def test_small_sphere_far_field_conditioning():
    """0.25 lambda leaf boxes (icosphere, r = lambda/2): the reference's K = 19
    sum is ill-conditioned there; GPU and restatement agree to the rounding
    noise of that sum, the near field to 1e-10."""
    pp, tt = meshes.icosphere(3, 0.5 * LAM)
    o = O.Oracle(pp, tt, LAM, 2.25, formulation="PMCHWT")
    o.tree()
    ia, ja, v = o.near_csr(nthreads=16)
    o.mlfmm_setup(16)
    g = SieTriMlfmm(LAM, 2.25, formulation="PMCHWT").precalculation_tri(pp, tt)
    _, _, vg = g.near_csr()
    for k in range(4):
        assert rel(vg[:, k], v[:, k]) < 1e-10
    rng = np.random.default_rng(7)
    x = rng.standard_normal(2 * o.N) + 1j * rng.standard_normal(2 * o.N)
    assert rel(g.calculate_vector_b_tri(x), o.matvec(x, 16)) < 1e-5


def test_single_edge_mesh_is_all_near():
    pp = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0.1]], float) * 1e-7
    tt = np.array([[1, 2, 3], [2, 4, 3]], np.int32)
    o = O.Oracle(pp, tt, LAM, 2.25)
    o.tree()
    o.near_csr()
    o.mlfmm_setup(1)
    g = SieTriMlfmm(LAM, 2.25).precalculation_tri(pp, tt)
    x = np.array([1.0 + 2j, -0.5j])
    assert rel(g.calculate_vector_b_tri(x), o.matvec(x)) < 1e-12
