"""Run-time tuning of the near-field preconditioned solver (spk_autotune,
spk_set_tuning; no counterpart in the reference) on the 5 lambda test patch.

- the tuner times the batched fp8 near product with 1, 2 and 4 row tiles
  per workgroup and five inner lengths, and keeps a measured choice;
- the three tile maps give the same product (to fp32 accumulation order);
- solves with the tuned values reach the tolerance in the FP64 true
  residual, and the solution equals the default solver's to the tolerance
  times the conditioning;
- an ensemble tuned on its first mesh gives the untuned ensemble's
  intensities to 1e-4;
- spk_set_tuning rejects values outside its domain.
"""
import numpy as np
import pytest

from specklesim_amd import DeviceBuffer, SieTriMlfmm, ensemble_intensity, meshes, plane_wave
from specklesim_amd._native import SpeckleError

pytestmark = pytest.mark.gpu
LAM = 633e-9


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def tuned():
    pp, tt = meshes.rough_surface(40, 40, 5.0, 5.0, 0.1, 0.5, 6, LAM)
    s = SieTriMlfmm(LAM, 2.25, formulation="ICTF", normal_convention=2,
                    precision="mixed").precalculation_tri(pp, tt)
    angles = [plane_wave(th, 0.0, 2, "TM") for th in np.radians([0, 20, 40, 60])]
    res = s.autotune(angles)
    print("tuning:", res)
    return pp, tt, s, angles, res


def test_tuner_measures_and_chooses(tuned):
    _, _, _, _, res = tuned
    assert res["tiles_per_group"] in (1, 2, 4)
    assert all(t > 0 for t in res["tile_us"])
    assert res["tile_us"][(1, 2, 4).index(res["tiles_per_group"])] == min(res["tile_us"])
    assert sorted(res["inner_ms"]) == [20, 25, 30, 36, 40]
    conv = {k: v for k, v in res["inner_ms"].items() if v > 0}
    assert res["inner_iterations"] == min(conv, key=conv.get)


def test_tile_maps_give_the_same_product(tuned):
    _, _, s, _, res = tuned
    rng = np.random.default_rng(2)
    X = (rng.standard_normal((16, 2 * s.N)) + 1j * rng.standard_normal((16, 2 * s.N)))
    Y = {}
    for tpw in (1, 2, 4):
        s.set_tuning(tpw, res["inner_iterations"])
        Y[tpw] = s.near_product_lowp(X, fmt="fp8", batched=True).astype(np.complex128)
    s.set_tuning(res["tiles_per_group"], res["inner_iterations"])
    for tpw in (2, 4):
        assert rel(Y[tpw], Y[1]) < 1e-6, tpw


def test_tuned_solves_converge(tuned):
    pp, tt, s, angles, _ = tuned
    s64 = SieTriMlfmm(LAM, 2.25, formulation="ICTF", normal_convention=2).precalculation_tri(pp, tt)
    nb = 2 * s.N * 16
    bs = [DeviceBuffer(nb) for _ in angles]
    xs = [DeviceBuffer(nb) for _ in angles]
    for (k, e), b in zip(angles, bs):
        s.get_vector_b_tri_device(k, e, b.ptr)
    infos = s.gmres_run_batch_device([b.ptr for b in bs], [x.ptr for x in xs], tolerance=1e-6,
                                     use_initial_guess=False, near_field=True)
    for (k, e), x, inf in zip(angles, xs, infos):
        assert inf["info"] == 0, inf
        b = s64.get_vector_b_tri(k, e)
        xv = x.download()
        assert rel(s64.calculate_vector_b_tri(xv), b) <= 1.01e-6
        x64, i64 = s64.gmres_run(b, tolerance=1e-6, max_iterations=650)
        assert rel(xv, x64) < 3e-5


def test_tuned_ensemble_matches_untuned():
    ms = [meshes.rough_surface(c, c, c / 10, c / 10, 0.1, 0.5, 20 + i, LAM)
          for i, c in enumerate((30, 40, 36))]
    angles = [plane_wave(th, 0.0, 2, "TM") for th in np.radians([0, 30])]
    th = np.radians(np.linspace(-60, 60, 13))
    pts = 1e3 * LAM * np.stack([np.sin(th), np.zeros_like(th), np.cos(th)], axis=1)
    I0, info0 = ensemble_intensity(LAM, 2.25, ms, angles, pts, workers=2)
    I1, info1 = ensemble_intensity(LAM, 2.25, ms, angles, pts, workers=2, inner_iterations="tune")
    assert all(i["info"] == 0 for row in info1 for i in row)
    assert rel(I1, I0) < 1e-4


def test_set_tuning_rejects_bad_values(tuned):
    _, _, s, _, res = tuned
    with pytest.raises(SpeckleError):
        s.set_tuning(3, 0)
    with pytest.raises(SpeckleError):
        s.set_tuning(1, 65)
    s.set_tuning(res["tiles_per_group"], res["inner_iterations"])
