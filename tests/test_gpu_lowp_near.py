"""The near-field preconditioner's low-precision products (no counterpart in
the reference) against the FP64 near-field product of the same operator.

The flexible GMRES of SPK_PRECOND_NEAR_FIELD only needs these products to
approximate the near field, so a wrong lane layout, row scale, component
order or swapped halves would still converge, with more outer steps.  This
file pins each product directly, per vector:
- complex64 blocks: to complex64 rounding (1e-5);
- bf16 blocks (8 mantissa bits): to 5e-3, and visibly coarser than complex64;
- fp8 E4M3FN blocks with per-row component scales (4 mantissa bits): to
  4e-2, and visibly coarser than bf16;
- each batched launch equals its single-vector counterpart to the rounding
  of x that the batched kernel adds (the matrix-core kernels stage x in
  bf16: 5e-3), for R = 1, 4, 9, 12 and 16 right-hand sides.
"""
import numpy as np
import pytest

from specklesim_amd import SieTriMlfmm, meshes

pytestmark = pytest.mark.gpu
LAM = 633e-9


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def case():
    pp, tt = meshes.rough_surface(40, 40, 5.0, 5.0, 0.1, 0.5, 2, LAM)
    s = SieTriMlfmm(LAM, 2.25, formulation="ICTF", normal_convention=2,
                    precision="mixed").precalculation_tri(pp, tt)
    rng = np.random.default_rng(21)
    X = (rng.standard_normal((16, 2 * s.N)) + 1j * rng.standard_normal((16, 2 * s.N))
         ).astype(np.complex64)
    Y = np.stack([s.matvec_parts(X[r].astype(np.complex128))[0] for r in range(16)])
    return s, X, Y


BOUNDS = {"c64": (0.0, 1e-5), "bf16": (1e-5, 5e-3), "fp8": (5e-4, 4e-2)}


@pytest.mark.parametrize("fmt", ["c64", "bf16", "fp8"])
def test_single_vector_products(case, fmt):
    s, X, Y = case
    Yl = s.near_product_lowp(X[:3], fmt, batched=False)
    lo, hi = BOUNDS[fmt]
    for r in range(3):
        e = rel(Yl[r].astype(np.complex128), Y[r])
        assert lo <= e < hi, (fmt, r, e)
        # both halves [xe; xh] separately (a swapped component order fails here)
        n = s.N
        assert rel(Yl[r][:n].astype(np.complex128), Y[r][:n]) < hi
        assert rel(Yl[r][n:].astype(np.complex128), Y[r][n:]) < hi


@pytest.mark.parametrize("fmt", ["bf16", "fp8"])
@pytest.mark.parametrize("R", [1, 4, 9, 12, 16])
def test_batched_products(case, fmt, R):
    s, X, Y = case
    Yb = s.near_product_lowp(X[:R], fmt, batched=True)
    # fp8 batches below 9 vectors run on the complex64 blocks
    ref_fmt = "c64" if (fmt == "fp8" and R < 9) else fmt
    Ys = s.near_product_lowp(X[:R], ref_fmt, batched=False)
    lo, hi = BOUNDS[ref_fmt]
    for r in range(R):
        assert rel(Yb[r], Ys[r]) < 5e-3, (fmt, R, r)
        assert rel(Yb[r].astype(np.complex128), Y[r]) < hi, (fmt, R, r)
    # the slots do not mix: permuting the batch permutes the outputs
    if R >= 9:
        perm = np.roll(np.arange(R), 3)
        Yp = s.near_product_lowp(X[:R][perm], fmt, batched=True)
        assert np.array_equal(Yp, Yb[perm])
