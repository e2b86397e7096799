"""Result-file formats of a GMRES run (specklesim_amd.results): the Fortran
ES19.12 edit, the Result_SEH_* and iteration_error_* files, and the output
names of main.f90.  The reference ships no output files, so the edit is
pinned by hand-checked strings of the Fortran standard's ESw.d rule."""
import numpy as np

from specklesim_amd import results as R


def test_fortran_es_edit():
    assert R.fortran_es(1.5) == " 1.500000000000E+00"
    assert R.fortran_es(-1.5) == "-1.500000000000E+00"
    assert R.fortran_es(0.0) == " 0.000000000000E+00"
    assert R.fortran_es(123456.789) == " 1.234567890000E+05"
    assert R.fortran_es(-2.5e-7) == "-2.500000000000E-07"
    # beyond two exponent digits the letter is dropped
    assert R.fortran_es(1e-100) == " 1.000000000000-100"
    assert R.fortran_es(-3e150) == "-3.000000000000+150"
    assert R._parse_fortran_real("1.000000000000-100") == 1e-100
    assert R._parse_fortran_real("-3.000000000000+150") == -3e150
    for v in (1.5, -2.5e-7, 1e-100, 0.1, 1 / 3):
        assert len(R.fortran_es(v)) == 19


def test_result_seh_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(14) + 1j * rng.standard_normal(14)) * 10.0 ** rng.integers(-8, 3, 14)
    x[3] = 0
    p = tmp_path / "Result_SEH_tri_PMCHWT.txt"
    R.write_result_seh(p, 11, x)
    lines = p.read_text().splitlines()
    assert lines[0].split() == ["n_el,", "11"] and lines[1].split() == ["m_pairs", "7"]
    assert len(lines) == 2 + 14
    assert all(len(l) == 19 + 5 + 19 for l in lines[2:])
    n_el, m, y = R.read_result_seh(p)
    assert (n_el, m) == (11, 7)
    # 13 significant digits survive the edit
    assert np.all(np.abs(y - x) <= 1e-12 * np.abs(x) + 1e-300)


def test_iteration_error_lines(tmp_path):
    hist = np.array([0.9, 0.3, 2e-3, 9e-4, 5e-4])
    checks = R.convergence_checks(hist, 1e-3, 650)
    assert checks == [(4, 9e-4), (5, 5e-4)]
    # an unconverged solve reports its last iteration
    assert R.convergence_checks(hist[:3], 1e-3, 3) == [(3, 2e-3)]
    p = tmp_path / "iteration_error_tri_NoP.txt"
    R.write_iteration_error(p, hist, 1e-3, 650)
    assert p.read_text().splitlines()[0] == "         4      9.000000000000E-04"
    its, bes = R.read_iteration_error(p)
    assert list(its) == [4, 5] and np.allclose(bes, [9e-4, 5e-4], rtol=1e-12)


def test_output_names():
    n = R.output_names("PMCHWT")
    assert n == {"result": "Result_SEH_tri_PMCHWT.txt", "efield": "Result_Efield_tri_BRDF_n.txt",
                 "iteration_error": "iteration_error_tri_NoP.txt"}
    assert R.output_names("ICTF", evaluation="cyl_field")["efield"] == "Result_Efield_tri_cyl_field."
    n = R.output_names("ICTF", 7, preconditioned=True)
    assert n["result"] == "Result_SEH_tri_ICTF_007.txt"
    assert n["iteration_error"] == "iteration_error_ICTF_007_LeftP.txt"
    assert R.output_names("MCTF", 12, multi_mesh=True)["iteration_error"] == \
        "iteration_error_MCTF.txt_NoP.txt"
    assert R.output_names("MCTF", 3, element="quad")["result"] == "Result_SEH_quad_MCTF_003.txt"
