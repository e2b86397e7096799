"""GPU quadrilateral fill (spk_quad_*) against the restatement in
oracle/spk_quad.c: integer numbering bit-exact; impedance_edge_versus_element
for every pair class (regular, corner-adjacent, edge-adjacent, self) and the
assembled systems of PMCHWT, MCTF and ICTF to 1e-10 relative (FP64; the GPU
differs by evaluation order and contraction only)."""
import numpy as np
import pytest

from oracle import oracle as O
from specklesim_amd import QuadFill, meshes

pytestmark = pytest.mark.gpu
LAM = 633e-9


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def grid():
    Lx, Ly, nx, ny, z = meshes.rough_height_field(3, 3, 1.2, 1.2, 0.05, 0.5, 2, LAM)
    q = O.QuadOracle(LAM, 2.25, Lx, Ly, nx, ny, z)
    g = QuadFill(LAM, 2.25, formulation="PMCHWT").setup(Lx, Ly, nx, ny, z)
    return q, g, (Lx, Ly, nx, ny, z)


def test_numbering(grid):
    q, g, _ = grid
    assert (g.M, g.n_el) == (q.M, q.n_el)
    assert np.array_equal(g.edges(), q.edges())
    pp, v = g.mesh()
    assert np.array_equal(pp, q.nodes()) and np.array_equal(v, q.elements())


@pytest.mark.parametrize("cls", ["self", "edge", "corner", "regular"])
def test_edge_element_classes(grid, cls):
    q, g, _ = grid
    e = q.edges()
    # element 5 is the centre of the 3 x 3 grid: 2, 4, 6, 8 share a side,
    # 1, 3, 7, 9 a corner; pairs on element 1 against 9 share nothing
    if cls == "self":
        pairs = [(s, e[s, 0]) for s in range(q.M) if e[s, 0] in (1, 5)]
    elif cls == "edge":
        pairs = [(s, t) for s in range(q.M) if e[s, 0] == 5 for t in (2, 4, 6, 8)]
    elif cls == "corner":
        pairs = [(s, t) for s in range(q.M) if e[s, 0] == 5 for t in (1, 3, 7, 9)]
    else:
        pairs = [(s, t) for s in range(q.M) if e[s, 0] == 1 for t in (6, 8, 9)]
    s = np.array([p[0] for p in pairs])
    t = np.array([p[1] - 1 for p in pairs])
    out = g.impedance_edge_versus_element(s, t)
    ref = np.array([q.edge_element(a + 1, b + 1) for a, b in zip(s, t)])
    for k in range(3):
        assert rel(out[:, k], ref[:, k]) < 1e-10, (cls, k)


@pytest.mark.parametrize("form", ["PMCHWT", "MCTF", "ICTF"])
def test_assembled_system(form):
    Lx, Ly, nx, ny, z = meshes.rough_height_field(3, 2, 1.0, 0.7, 0.05, 0.5, 5, LAM)
    eps2 = 18.392 - 0.4164j
    q = O.QuadOracle(LAM, eps2, Lx, Ly, nx, ny, z)
    g = QuadFill(LAM, eps2, formulation=form).setup(Lx, Ly, nx, ny, z)
    MM = g.fill()
    ref = q.assemble(form, nthreads=16)
    assert rel(MM, ref) < 1e-10
    # every block row to 1e-9 of its own scale
    scale = np.abs(ref).max(axis=1)
    assert np.max(np.abs(MM - ref).max(axis=1) / scale) < 1e-9
    assert np.array_equal(g.fill(), MM)


@pytest.mark.parametrize("form", ["PMCHWT", "MCTF", "ICTF"])
def test_plane_wave_rhs(form):
    """get_vector_b_quad (src/lib_sie_quad_calculation_mod.f90:91-138) on the
    GPU against the restatement, lossy exterior and interior, oblique
    incidence: 1e-12 relative."""
    Lx, Ly, nx, ny, z = meshes.rough_height_field(4, 3, 1.5, 1.1, 0.05, 0.5, 9, LAM)
    eps1, eps2 = 1.3 - 0.05j, 18.392 - 0.4164j
    q = O.QuadOracle(LAM, eps2, Lx, Ly, nx, ny, z, eps_r1=eps1)
    g = QuadFill(LAM, eps2, formulation=form, eps_r1=eps1).setup(Lx, Ly, nx, ny, z)
    th = np.deg2rad(25.0)
    k = np.array([np.sin(th), 0.0, -np.cos(th)])
    for e in (np.array([0.0, 1.0, 0.0]), np.array([np.cos(th), 0.0, np.sin(th)])):
        b = g.get_vector_b_quad(k, e)
        ref = q.rhs_plane(k, e, form)
        M = q.M
        assert rel(b[:M], ref[:M]) < 1e-12 and rel(b[M:], ref[M:]) < 1e-12


@pytest.mark.parametrize("form", ["PMCHWT", "MCTF", "ICTF"])
def test_lu_solve_and_field(form):
    """lib_sie_quad_solver_normal_I/II (src/lib_sie_quad_solver_normal.f90:80-335):
    GPU fill, right-hand side and rocSOLVER LU against LAPACK's ZGESV (the
    routine solver_LU calls, src/lib_sie_math.f90:649-676) on the oracle's
    system and right-hand side; then the field of the solution against the
    oracle's Integration_scattered_field_quad."""
    Lx, Ly, nx, ny, z = meshes.rough_height_field(5, 4, 1.8, 1.4, 0.05, 0.5, 11, LAM)
    eps2 = 18.392 - 0.4164j
    q = O.QuadOracle(LAM, eps2, Lx, Ly, nx, ny, z)
    g = QuadFill(LAM, eps2, formulation=form).setup(Lx, Ly, nx, ny, z)
    th = np.deg2rad(30.0)
    k = np.array([np.sin(th), 0.0, -np.cos(th)])
    e = np.array([0.0, 1.0, 0.0])
    x = g.solve_plane_wave(k, e)
    A = q.assemble(form, nthreads=16)
    b = q.rhs_plane(k, e, form)
    xr = np.linalg.solve(A, b)
    cond = np.linalg.cond(A)
    res = rel(A @ x, b)
    ex = rel(x, xr)
    print(f"{form}: 2M {2 * g.M}, cond {cond:.1e}, residual {res:.1e}, vs ZGESV {ex:.1e}")
    # backward stable under the oracle's system; forward error within the
    # conditioning of the system
    assert res < 1e-12
    assert ex < max(1e-10, 1e-15 * cond)
    pts = np.array([[0.3, -0.2, 1.0], [1.5, 0.4, 2.0], [-2.0, 1.0, 0.5]]) * LAM * 3
    E, H = g.scattered_field(pts, x)
    Eo, Ho = q.scattered_field(pts, x)
    assert rel(E, Eo) < 1e-11 and rel(H, Ho) < 1e-11


def test_lu_solve_several_right_hand_sides_and_refusals():
    """Two right-hand sides in one solve equal two single solves; a matrix
    with a zero column is refused as singular, as the reference stops."""
    from specklesim_amd import _native as nat
    Lx, Ly, nx, ny, z = meshes.rough_height_field(3, 3, 1.2, 1.2, 0.05, 0.5, 4, LAM)
    g = QuadFill(LAM, 2.25, formulation="ICTF").setup(Lx, Ly, nx, ny, z)
    n = 2 * g.M
    ks = [np.array([0.0, 0.0, -1.0]), np.array([np.sin(0.4), 0.0, -np.cos(0.4)])]
    es = [np.array([1.0, 0.0, 0.0]), np.array([0.0, 1.0, 0.0])]
    single = [g.solve_plane_wave(k, e) for k, e in zip(ks, es)]
    MM = nat.DeviceBuffer(n * n * 16)
    B = nat.DeviceBuffer(2 * n * 16)
    g.fill_device(MM.ptr)
    for j in range(2):
        g.get_vector_b_quad_device(ks[j], es[j], B.ptr + j * n * 16)
    g.solve_lu_device(MM.ptr, B.ptr, nrhs=2)
    X = B.download(np.complex128).reshape(2, n)
    for j in range(2):
        assert rel(X[j], single[j]) < 1e-13
    A = g.fill()
    A[:, 5] = 0
    MM.upload(A)
    with pytest.raises(nat.SpeckleError, match="singular"):
        g.solve_lu_device(MM.ptr, B.ptr, nrhs=1)
