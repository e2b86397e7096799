"""Flexible GMRES on the FP64 operator, right-preconditioned by a short
GMRES solve on the near-field operator (SPK_PRECOND_NEAR_FIELD): it reaches
the FP64 tolerance in the true residual with a fraction of the outer steps
of plain GMRES, and its solution agrees with plain GMRES's to the tolerance
(both solve the same FP64 system; the difference is bounded by the
conditioning times 1e-6, measured 3.7e-6 here), on FP64 handles (FP64 inner
solves) and mixed handles (complex64 inner solves), one right-hand side or a
batch."""
import numpy as np
import pytest

from specklesim_amd import DeviceBuffer, SieTriMlfmm, farfield, meshes, plane_wave

pytestmark = pytest.mark.gpu
LAM = 633e-9


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def mesh():
    return meshes.rough_surface(40, 40, 5.0, 5.0, 0.1, 0.5, 1, LAM)


@pytest.fixture(scope="module")
def solvers(mesh):
    pp, tt = mesh
    kw = dict(formulation="ICTF", normal_convention=2, tree_s_opt=250, truncation_number=19)
    return {p: SieTriMlfmm(LAM, 2.25, precision=p, **kw).precalculation_tri(pp, tt)
            for p in ("fp64", "mixed")}


@pytest.fixture(scope="module")
def reference(solvers):
    """Plain FP64 GMRES (IMGS) solutions for three incidence angles."""
    s = solvers["fp64"]
    out = []
    for a in (0.0, 20.0, 40.0):
        k, e = plane_wave(np.deg2rad(a), 0.0, 2, "TM")
        b = s.get_vector_b_tri(k, e)
        x, inf = s.gmres_run(b, tolerance=1e-6, max_iterations=650, use_initial_guess=False)
        assert inf["info"] == 0
        out.append((a, b, x, inf["iters"]))
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed", "mixed-bf16"])
def test_near_field_fgmres_single(solvers, reference, precision, monkeypatch):
    """mixed: the inner products of one right-hand side run on the fp8
    (E4M3FN, per-row component scales) copy of the near field; mixed-bf16:
    on the bf16 copy (SPK_NEAR_F8=0), as batches of four or more do."""
    if precision == "mixed-bf16":
        monkeypatch.setenv("SPK_NEAR_F8", "0")
    s = solvers[precision.split("-")[0]]
    a, b, x0, it0 = reference[0]
    x, inf = s.gmres_run(b, tolerance=1e-6, max_iterations=650, use_initial_guess=False,
                         near_field=True)
    assert inf["info"] == 0 and inf["be"] <= 1e-6, inf
    res = np.linalg.norm(b - solvers["fp64"].calculate_vector_b_tri(x)) / np.linalg.norm(b)
    assert res <= 1e-6 and abs(res - inf["be"]) < 1e-9, (res, inf["be"])
    assert inf["iters"] * 4 < it0, (inf["iters"], it0)
    assert inf["matvecs"] == inf["iters"] + 1
    assert rel(x, x0) < 3e-5
    _, I0 = farfield.brdf(solvers["fp64"], x0, 0.0, (-1.2, 1.2), 41, 0.0, 1e3 * LAM)
    _, I1 = farfield.brdf(solvers["fp64"], x, 0.0, (-1.2, 1.2), 41, 0.0, 1e3 * LAM)
    assert np.max(np.abs(I1 - I0)) / np.max(np.abs(I0)) < 1e-4
    hist = s.gmres_history()
    assert len(hist) == inf["iters"] and hist[-1] <= 1e-6 and np.all(np.diff(hist) <= 1e-12)


def test_near_field_fgmres_fp8_against_bf16(solvers, reference, monkeypatch):
    """The fp8 copy of the near field (one right-hand side) preconditions
    about as well as the bf16 one: at most two more outer steps, and both
    solutions agree with plain GMRES's."""
    s = solvers["mixed"]
    _, b, x0, _ = reference[2]
    x8, i8 = s.gmres_run(b, tolerance=1e-6, max_iterations=650, use_initial_guess=False,
                         near_field=True)
    monkeypatch.setenv("SPK_NEAR_F8", "0")
    xb, ib = s.gmres_run(b, tolerance=1e-6, max_iterations=650, use_initial_guess=False,
                         near_field=True)
    assert i8["info"] == 0 and ib["info"] == 0
    assert i8["iters"] <= ib["iters"] + 2, (i8["iters"], ib["iters"])
    assert rel(x8, x0) < 3e-5 and rel(xb, x0) < 3e-5


def test_near_field_fgmres_relaxed_products(solvers, reference, monkeypatch):
    """On a mixed handle the outer products switch to the complex64 operator
    once the residual estimate is at or below 2e-4 (tolerance 1e-6): the
    solve still meets the tolerance in the FP64 true residual, in as many
    outer steps as with every product in FP64 (SPK_RELAX=0), and both agree
    with plain GMRES."""
    s = solvers["mixed"]
    _, b, x0, _ = reference[1]
    xr, ir = s.gmres_run(b, tolerance=1e-6, max_iterations=650, use_initial_guess=False,
                         near_field=True)
    monkeypatch.setenv("SPK_RELAX", "0")
    xf, iff = s.gmres_run(b, tolerance=1e-6, max_iterations=650, use_initial_guess=False,
                          near_field=True)
    assert ir["info"] == 0 and iff["info"] == 0
    res = np.linalg.norm(b - solvers["fp64"].calculate_vector_b_tri(xr)) / np.linalg.norm(b)
    assert res <= 1e-6 and abs(res - ir["be"]) < 1e-9, (res, ir["be"])
    assert ir["iters"] <= iff["iters"] + 1, (ir["iters"], iff["iters"])
    assert rel(xr, x0) < 3e-5 and rel(xf, x0) < 3e-5


@pytest.mark.parametrize("inner", [1, 8, 100])
def test_near_field_fgmres_inner_lengths(solvers, reference, inner):
    """Any inner length converges (100 is clamped to 64); one inner step is
    already a preconditioner (the first Krylov direction scaled)."""
    s = solvers["mixed"]
    _, b, x0, it0 = reference[1]
    x, inf = s.gmres_run(b, tolerance=1e-6, max_iterations=650, use_initial_guess=False,
                         near_field=True, inner_iterations=inner)
    assert inf["info"] == 0 and inf["be"] <= 1e-6
    assert rel(x, x0) < 3e-5


def test_near_field_fgmres_restart_and_initial_guess(solvers, reference):
    """Short cycles (restart 5) still converge from the true residual of
    each cycle, and a converged initial guess returns after one residual."""
    s = solvers["mixed"]
    _, b, x0, _ = reference[2]
    x, inf = s.gmres_run(b, tolerance=1e-6, max_iterations=650, restart=5,
                         use_initial_guess=False, near_field=True)
    assert inf["info"] == 0 and inf["be"] <= 1e-6 and rel(x, x0) < 3e-5
    x2, inf2 = s.gmres_run(b, x0=x, tolerance=1e-6, max_iterations=650, use_initial_guess=True,
                           near_field=True)
    assert inf2["iters"] == 0 and inf2["matvecs"] == 1 and np.array_equal(x2, x)


@pytest.mark.parametrize("precision,R", [("mixed", 2), ("mixed", 6), ("mixed", 12),
                                         ("mixed-bf16", 12), ("fp64", 5)])
def test_near_field_fgmres_batch(solvers, reference, precision, R, monkeypatch):
    """R right-hand sides together (batched products from 4 up) reach the
    tolerance each in its own true residual; mixed with 9 or more: the
    matrix-core preconditioner product on the fp8 copy (mixed-bf16: on the
    bf16 copy, SPK_NEAR_F8=0)."""
    if precision == "mixed-bf16":
        monkeypatch.setenv("SPK_NEAR_F8", "0")
    s = solvers[precision.split("-")[0]]
    nb = 2 * s.N * 16
    angles = np.linspace(0.0, 50.0, R)
    bs, xs, rhs = [], [], []
    for a in angles:
        k, e = plane_wave(np.deg2rad(a), 0.0, 2, "TM")
        b = s.get_vector_b_tri(k, e)
        rhs.append(b)
        bs.append(DeviceBuffer(nb).upload(b))
        xs.append(DeviceBuffer(nb))
    infos = s.gmres_run_batch_device([d.ptr for d in bs], [d.ptr for d in xs], max_iterations=650,
                                     tolerance=1e-6, use_initial_guess=False, near_field=True)
    for r, (b, inf) in enumerate(zip(rhs, infos)):
        assert inf["info"] == 0 and inf["be"] <= 1e-6, (r, inf)
        x = xs[r].download()
        res = np.linalg.norm(b - solvers["fp64"].calculate_vector_b_tri(x)) / np.linalg.norm(b)
        assert res <= 1e-6 and abs(res - inf["be"]) < 1e-9, (r, res, inf["be"])
    _, b0, x0, _ = reference[0]
    assert rel(xs[0].download(), x0) < 3e-5


@pytest.mark.parametrize("R", [1, 5, 12])
def test_inner_solve_graph_replays_equal_eager(solvers, R, monkeypatch):
    """The inner near-field solve is captured once as a HIP graph and
    replayed at every outer step (19 replays a solve, with the outer step's
    eager launches and host synchronisations between them): the solutions
    are bit-identical to launching the inner steps one by one, on two solves
    in a row (the second reuses the captured graph) and after the handle is
    set up again for another mesh and back (the graph is rebuilt)."""
    s = solvers["mixed"]
    angles = np.linspace(0.0, 50.0, R)

    def solve():
        nb = 2 * s.N * 16
        bs, xs = [], []
        for a in angles:
            k, e = plane_wave(np.deg2rad(a), 0.0, 2, "TM")
            bs.append(DeviceBuffer(nb).upload(s.get_vector_b_tri(k, e)))
            xs.append(DeviceBuffer(nb))
        infos = s.gmres_run_batch_device([d.ptr for d in bs], [d.ptr for d in xs],
                                         max_iterations=650, tolerance=1e-6,
                                         use_initial_guess=False, near_field=True)
        return [x.download() for x in xs], [i["iters"] for i in infos]

    monkeypatch.setenv("SPK_INNER_GRAPH", "0")
    x_eager, it_eager = solve()
    monkeypatch.delenv("SPK_INNER_GRAPH")
    for _ in range(2):
        x_g, it_g = solve()
        assert it_g == it_eager
        for a, b in zip(x_g, x_eager):
            assert np.array_equal(a, b)
    pp2, tt2 = meshes.rough_surface(30, 30, 3.0, 3.0, 0.1, 0.5, 2, LAM)
    s.precalculation_tri(pp2, tt2)
    _, it_other = solve()
    assert all(i < 60 for i in it_other)
    pp, tt = meshes.rough_surface(40, 40, 5.0, 5.0, 0.1, 0.5, 1, LAM)
    s.precalculation_tri(pp, tt)
    x_back, it_back = solve()
    assert it_back == it_eager
    for a, b in zip(x_back, x_eager):
        assert np.array_equal(a, b)
