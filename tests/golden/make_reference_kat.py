"""Extract the known-answer values that the reference's own self-tests hold
into tests/golden/reference_kat.json.

Run in the build container (reads the reference sources as text):
    python tests/golden/make_reference_kat.py /root/reference

Sources (reference repository paths):
  * spherical Hankel h_n^(2)(2+2i), n=0..40:
    src/lib/libmath/src/math/lib_math_bessel.f90:3045-3111
    (test_lib_math_bessel_spherical_third_kind_2_cmplx, sageMath values)
  * Legendre P_l(0.4), P'_l(0.4), l=0..5:
    src/lib/libmath/src/math/lib_math_legendre.f90:1918-1968
  * octree helpers (3-D branches):
    src/lib/libtree/src/lib/lib_tree_helper_functions.f90:2207-2700
  * l_max of five diagonal points for s = 1,2,3:
    src/lib/libtree/src/lib/lib_tree.f90:2970-3056
  * GMRES test problem (tridiagonal 4 / -2+i, x = ones):
    src/lib/libmath/src/math/lib_math_solver_GMRES.f90:381-500
"""
from __future__ import annotations

import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NUM = r"([-+]?\s*[0-9.]+(?:[dDeE][-+]?[0-9]+)?(?:_8)?)"


def _f(tok: str) -> float:
    tok = tok.replace(" ", "").replace("_8", "").replace("d", "e").replace("D", "e")
    return float(tok)


def hankel(ref: str):
    path = os.path.join(ref, "src/lib/libmath/src/math/lib_math_bessel.f90")
    lines = open(path).read().splitlines()
    seg = "\n".join(lines[3044:3112])
    out = {}
    for m in re.finditer(r"ground_truth_y\((\d+)\)\s*=\s*cmplx\(" + NUM + r"\s*,\s*" + NUM, seg):
        out[int(m.group(1))] = (_f(m.group(2)), _f(m.group(3)))
    n = max(out)
    return {"z": [2.0, 2.0], "source": "src/lib/libmath/src/math/lib_math_bessel.f90:3045-3111",
            "re": [out[i][0] for i in range(1, n + 1)], "im": [out[i][1] for i in range(1, n + 1)]}


def legendre(ref: str):
    path = os.path.join(ref, "src/lib/libmath/src/math/lib_math_legendre.f90")
    lines = open(path).read().splitlines()
    seg = "\n".join(lines[1917:1970])
    pm, pd = {}, {}
    for m in re.finditer(r"ground_truth_pm\((\d+)\)\s*=\s*" + NUM, seg):
        pm[int(m.group(1))] = _f(m.group(2))
    for m in re.finditer(r"ground_truth_pd\((\d+)\)\s*=\s*" + NUM, seg):
        pd[int(m.group(1))] = _f(m.group(2))
    return {"x": 0.4, "source": "src/lib/libmath/src/math/lib_math_legendre.f90:1918-1968",
            "pm": [pm[i] for i in sorted(pm)], "pd": [pd[i] for i in sorted(pd)]}


def tree():
    # 3-D branches of the reference's tree self-tests, transcribed as data.
    xs_lmax = [[float(np.float32(1.0) - np.float32(0.999 * i) / np.float32(5.0))] * 3
               for i in range(1, 6)]
    # the reference evaluates 1.0 - (0.999*i)/(1.0*5) in default real
    xs_lmax = [[float(np.float32(np.float32(1.0) - np.float32(np.float32(0.999) * np.float32(i))
                                 / np.float32(5.0)))] * 3 for i in range(1, 6)]
    return {
        "uindex": {"x": [0.75, 0.5, float(np.float32(2.0 ** -9 + 2.0 ** -8))], "l": 1, "n": 6,
                   "source": "lib_tree_helper_functions.f90:2207-2243"},
        "uindex_lmax": {"x": [1.0 - 2.0 ** -21] * 3, "l": 21, "n": 2 ** 63 - 1,
                        "source": "lib_tree_helper_functions.f90:2285-2297 (double branch)"},
        "parent": {"n": 6, "l": 1, "parent": 0, "source": "lib_tree_helper_functions.f90:2318-2351"},
        "children": {"n": 1, "children": list(range(8, 16)),
                     "source": "lib_tree_helper_functions.f90:2353-2397"},
        "centre": {"n": 1, "l": 1, "x": [0.25, 0.25, 0.75],
                   "source": "lib_tree_helper_functions.f90:2399-2438"},
        "neighbours": [
            {"k": 1, "l": 2, "n": 35,
             "list": [4, 5, 12, 6, 7, 14, 20, 21, 28, 32, 33, 40, 34, 42, 48, 49, 56, 36, 37, 44,
                      38, 39, 46, 52, 53, 60],
             "source": "lib_tree_helper_functions.f90:2440-2529"},
            {"k": 1, "l": 1, "n": 5, "list": [0, 1, 2, 3, 4, 6, 7],
             "source": "lib_tree_helper_functions.f90:2531-2612"}],
        "coordinate_binary": {"x": [0.25, 0.375, 0.4375],
                              "bits64": [2 ** 62, 2 ** 62 + 2 ** 61, 2 ** 62 + 2 ** 61 + 2 ** 60],
                              "source": "lib_tree_helper_functions.f90:2708-2764"},
        "level_min": {"k": 1, "l_min": 2, "source": "lib_tree.f90:2946-2968"},
        "level_max": {"xs": xs_lmax, "l_th": 16, "s": [1, 2, 3], "l_max": [3, 2, 1],
                      "sorted_order_1based": [5, 4, 3, 2, 1],
                      "source": "lib_tree.f90:2970-3056"},
    }


def gmres():
    return {"n": 100, "diag": [4.0, 0.0], "offdiag": [-2.0, 1.0], "restart": 60,
            "solution": "ones", "source": "src/lib/libmath/src/math/lib_math_solver_GMRES.f90:381-500"}


def main(ref: str):
    data = {"hankel2_sph": hankel(ref), "legendre": legendre(ref), "tree": tree(),
            "gmres": gmres()}
    with open(os.path.join(HERE, "reference_kat.json"), "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", os.path.join(HERE, "reference_kat.json"))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
