"""The oracle's Lagrange interpolation between sample grids
(Lagrange_Interpolation_New, src/lib_sie_mlfmm.f90:268-493) and its MLFMM
with per-level truncation numbers, pinned independently:

- interior stencils: on samples of a tensor polynomial of degree 5 in theta
  and in phi, every target whose 6 x 6 stencil stays inside the grid is
  reproduced to 1e-12 (the weights' nodes and the data rows and columns
  line up);
- the pole rows, reflections across the poles and the periodic extension:
  the theta / phi components of smooth vector fields (with their Cartesian
  pole samples) are interpolated to another grid with an error that falls
  with the grid (to 2e-5 at K = 16 -> 21 and 21 -> 16), up to the poles; the
  pole samples are copied exactly;
- the MLFMM: with truncation numbers sized to the boxes, the transpose of
  the upward interpolation in the downward pass (anterpolation, applied to
  the folded coefficients) keeps the operator within 2e-3 of the dense one,
  while the reference's own downward step (interpolating the local
  coefficients) is off by more than 10 %.
"""
import numpy as np
import pytest

from oracle import oracle as O
from specklesim_amd import meshes

LAM = 633e-9


def _field(K, vs, d, k=2 * np.pi):
    kh, tt, tp, _ = O.sample_grid(K)
    ph = np.exp(1j * k * kh @ d)
    vals = np.zeros((4, kh.shape[0]), complex)
    for med, v in enumerate(vs):
        vals[med] = (tt @ v) * ph
        vals[2 + med] = (tp @ v) * ph
    return vals


def _stencil_starts(Km, Kn):
    """The reference's stencil starts (float arithmetic of default reals)."""
    t = [int(np.floor(np.float32(np.float32(it - 1.0) / np.float32(Kn)) * np.float32(Km))) + 1
         for it in range(1, Kn + 1)]
    s = [int(np.floor(np.float32(np.float32(jp - 1.0) / np.float32(Kn)) * np.float32(Km))) + 1
         for jp in range(1, 2 * Kn + 1)]
    return np.array(t), np.array(s)


@pytest.mark.parametrize("Km,Kn", [(12, 17), (16, 12), (19, 20)])
def test_interior_stencils_reproduce_polynomials(Km, Kn):
    p = 3
    rng = np.random.default_rng(Km * 100 + Kn)
    a, b = rng.standard_normal(6), rng.standard_normal(6)
    P = lambda th: np.polyval(a, th)
    Q = lambda ph: np.polyval(b, ph)
    xm, _ = np.polynomial.legendre.leggauss(Km)
    th_m, ph_m = np.arccos(xm), np.arange(2 * Km) * np.pi / Km
    vals = np.zeros((4, 2 * Km * Km + 2), complex)
    for q in range(4):
        vals[q, :-2] = (np.outer(P(th_m), Q(ph_m)) * (1 + 0.5j * q)).ravel()
    out = O.lagrange_interp(vals, Km, Kn, p)
    xn, _ = np.polynomial.legendre.leggauss(Kn)
    th_n, ph_n = np.arccos(xn), np.arange(2 * Kn) * np.pi / Kn
    t, s = _stencil_starts(Km, Kn)
    # ext rows t .. t+2p-1 (0-based) hold interior nodes p+1 .. p+Km; ext
    # columns s .. s+2p-1 the interior phi indices p .. p+2Km-1
    rows = np.nonzero((t >= p + 1) & (t + 2 * p - 1 <= p + Km))[0]
    cols = np.nonzero((s >= p) & (s + 2 * p - 1 <= p + 2 * Km - 1))[0]
    assert rows.size >= Kn // 2 and cols.size >= Kn
    scale = np.abs(vals).max()
    for q in range(4):
        got = out[q, :-2].reshape(Kn, 2 * Kn)[np.ix_(rows, cols)]
        ref = np.outer(P(th_n[rows]), Q(ph_n[cols])) * (1 + 0.5j * q)
        assert np.abs(got - ref).max() < 1e-12 * scale, q


def test_smooth_fields_converge_up_to_the_poles():
    vs = [np.array([0.3, -0.7, 0.5]), np.array([1.0, 0.2, -0.4])]
    d = np.array([0.2, 0.15, -0.1])
    errs = []
    for Km, Kn in ((8, 13), (12, 17), (16, 21), (21, 16)):
        src = _field(Km, vs, d)
        out = O.lagrange_interp(src, Km, Kn)
        ref = _field(Kn, vs, d)
        assert np.array_equal(out[:, -2:], src[:, -2:])  # pole samples copied
        e = np.abs(out - ref).max() / np.abs(ref).max()
        # the worst error anywhere, the first and last theta rows included
        errs.append(e)
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 2e-4 and errs[3] < 2e-4, errs


@pytest.fixture(scope="module")
def patch3():
    pp, tt = meshes.rough_surface(30, 30, 3.0, 3.0, 0.1, 0.5, 2, LAM)
    o = O.Oracle(pp, tt, LAM, 2.25, formulation="ICTF", normal_convention=2, s_opt=60)
    T = o.tree()
    assert (T["l_min"], T["l_max"]) == (2, 3)
    o.near_csr(nthreads=8)
    o.mlfmm_setup(8)
    rng = np.random.default_rng(7)
    x = rng.standard_normal(2 * o.N) + 1j * rng.standard_normal(2 * o.N)
    return pp, tt, x, o.matvec(x, 8), o.dense(8) @ x


def _mk(pp, tt, level_K, rr):
    o = O.Oracle(pp, tt, LAM, 2.25, formulation="ICTF", normal_convention=2, s_opt=60)
    o.tree()
    o.near_csr(nthreads=8)
    o.mlfmm_setup(8, level_K=level_K, rr=rr)
    return o


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_per_level_K_matvec_against_dense(patch3):
    pp, tt, x, y19, yd = patch3
    assert rel(y19, yd) < 1e-4
    # equal truncation numbers through level_K: the same operator
    assert np.array_equal(_mk(pp, tt, [0, 0, 19, 19], "adjoint").matvec(x, 8), y19)
    ya = _mk(pp, tt, [0, 0, 16, 12], "adjoint").matvec(x, 8)
    yu12 = _mk(pp, tt, [0, 0, 12, 12], "adjoint").matvec(x, 8)
    yr = _mk(pp, tt, [0, 0, 16, 12], "reference").matvec(x, 8)
    ea, e12, er = rel(ya, yd), rel(yu12, yd), rel(yr, yd)
    print(f"K 16/12 anterpolation {ea:.2e}, uniform 12 {e12:.2e}, reference RR {er:.2e}")
    assert ea < 2e-3 and ea < e12
    assert er > 0.1


def test_level_K_must_not_grow_towards_the_leaves(patch3):
    pp, tt, *_ = patch3
    with pytest.raises(ValueError):
        _mk(pp, tt, [0, 0, 12, 16], "adjoint")


@pytest.mark.parametrize("Km,Kn", [(12, 17), (16, 12), (19, 20), (21, 16)])
def test_library_interpolation_equals_the_oracle(Km, Kn):
    """The library's construction (a symbolic replay of the extended grid,
    applied as a sparse matrix; host probe, no GPU) against the oracle's
    literal restatement, on random samples."""
    import ctypes as C

    from specklesim_amd import _native as nat

    rng = np.random.default_rng(Km + 7 * Kn)
    vals = rng.standard_normal((4, 2 * Km * Km + 2)) + 1j * rng.standard_normal((4, 2 * Km * Km + 2))
    ref = O.lagrange_interp(vals, Km, Kn)
    out = np.zeros((4, 2 * Kn * Kn + 2), complex)
    dp = C.POINTER(C.c_double)
    vin = np.ascontiguousarray(vals)
    nat.check(nat.lib().spk_interp_probe(Km, Kn, vin.view(np.float64).ctypes.data_as(dp),
                                         out.view(np.float64).ctypes.data_as(dp)))
    assert np.abs(out - ref).max() < 1e-13 * np.abs(vals).max() * 36
