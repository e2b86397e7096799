"""8-node quadrilateral (10 Huber edge functions) near/dense fill on the GPU.

Mirrors the reference's quad path (src/lib_sie_quad_solver_normal.f90:80-210):
``setup`` builds the height-field mesh, neighbours and edge numbering
(sorting_quad, Neighbours_EdgeCorner_Full, Sorting_edge);
``impedance_edge_versus_element`` is the reference routine of that name
(src/lib_sie_quad_singularity.f90:22-98) for (edge function, element) pairs,
0-based here; ``fill`` returns the assembled 2M x 2M system,
``solve_plane_wave`` solves it by LU as the reference does and
``scattered_field`` evaluates the field of a solution.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as nat
from .sie import FORMULATIONS


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class QuadFill:
    def __init__(self, wavelength, eps_r2, eps_r1=1.0 + 0.0j, formulation="PMCHWT", device=0):
        L = nat.lib()
        f = FORMULATIONS[formulation] if isinstance(formulation, str) else int(formulation)
        e1, e2 = complex(eps_r1), complex(eps_r2)
        self.params = nat.Params(wavelength, e1.real, e1.imag, e2.real, e2.imag, f, 1, 3, 250, 19,
                                 0, device)
        h = C.c_void_p()
        nat.check(L.spk_quad_create(C.byref(self.params), C.byref(h)))
        self._h = h
        self.M = self.n_el = self.n_nodes = 0

    def close(self):
        if getattr(self, "_h", None):
            nat.lib().spk_quad_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def setup(self, Lx, Ly, nx, ny, z):
        z = np.ascontiguousarray(z, dtype=np.float64).reshape(-1)
        if z.size != (2 * nx + 1) * (2 * ny + 1):
            raise ValueError("z must hold (2nx+1)(2ny+1) heights")
        nat.check(nat.lib().spk_quad_setup(self._h, Lx, Ly, nx, ny, _dp(z)))
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        nat.check(nat.lib().spk_quad_sizes(self._h, C.byref(a), C.byref(b), C.byref(c)))
        self.M, self.n_el, self.n_nodes = a.value, b.value, c.value
        return self

    def mesh(self):
        pp = np.zeros((self.n_nodes, 3))
        v = np.zeros((self.n_el, 8), np.int32)
        nat.check(nat.lib().spk_quad_get_mesh(self._h, _dp(pp), _ip(v)))
        return pp, v

    def edges(self):
        e = np.zeros((self.M, 5), np.int32)
        nat.check(nat.lib().spk_quad_get_edges(self._h, _ip(e)))
        return e

    def impedance_edge_versus_element(self, s, t):
        s = np.ascontiguousarray(np.atleast_1d(s), dtype=np.int32)
        t = np.ascontiguousarray(np.atleast_1d(t), dtype=np.int32)
        out = np.zeros(60 * s.size)
        nat.check(nat.lib().spk_quad_edge_element(self._h, s.size, _ip(s), _ip(t), _dp(out)))
        return (out[0::2] + 1j * out[1::2]).reshape(-1, 3, 10)

    def fill(self):
        MM = np.zeros((2 * self.M, 2 * self.M), np.complex128)
        nat.check(nat.lib().spk_quad_fill(self._h, _dp(MM.view(np.float64))))
        return MM

    def fill_device(self, d_MM):
        nat.check(nat.lib().spk_quad_fill_device(self._h, C.c_void_p(int(d_MM))))

    def get_vector_b_quad(self, k_in, e_in):
        """Plane-wave right-hand side (get_vector_b_quad,
        src/lib_sie_quad_calculation_mod.f90:91-138): b[2M] complex."""
        k = np.ascontiguousarray(k_in, dtype=np.float64)
        e = np.ascontiguousarray(e_in, dtype=np.float64)
        b = np.zeros(2 * self.M, np.complex128)
        nat.check(nat.lib().spk_quad_rhs_plane_wave(self._h, _dp(k), _dp(e),
                                                    _dp(b.view(np.float64))))
        return b

    def get_vector_b_quad_device(self, k_in, e_in, d_b):
        k = np.ascontiguousarray(k_in, dtype=np.float64)
        e = np.ascontiguousarray(e_in, dtype=np.float64)
        nat.check(nat.lib().spk_quad_rhs_plane_wave_device(self._h, _dp(k), _dp(e),
                                                           C.c_void_p(int(d_b))))

    def solve_lu_device(self, d_MM, d_b, nrhs=1):
        """solver_LU (ZGESV, src/lib_sie_math.f90:649-676) on the device:
        d_MM (row-major, from fill_device) becomes its LU factors, d_b
        (nrhs columns of 2M) the solutions."""
        nat.check(nat.lib().spk_quad_solve_lu_device(self._h, C.c_void_p(int(d_MM)), int(nrhs),
                                                     C.c_void_p(int(d_b))))

    def solve_plane_wave(self, k_in, e_in):
        """lib_sie_quad_solver_normal_I (src/lib_sie_quad_solver_normal.f90:80-229):
        fill, plane-wave right-hand side and LU solve; V_EH[2M] complex."""
        k = np.ascontiguousarray(k_in, dtype=np.float64)
        e = np.ascontiguousarray(e_in, dtype=np.float64)
        x = np.zeros(2 * self.M, np.complex128)
        nat.check(nat.lib().spk_quad_solve_plane_wave(self._h, _dp(k), _dp(e),
                                                      _dp(x.view(np.float64))))
        return x

    def scattered_field(self, points, x):
        """lib_sie_quad_solver_normal_II (src/lib_sie_quad_solver_normal.f90:272-335):
        E, H (npts x 3 complex) of the solution x[2M] at the points."""
        x = np.ascontiguousarray(x, dtype=np.complex128)
        if x.size != 2 * self.M:
            raise ValueError("x must hold 2M coefficients")
        r = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        E = np.zeros((r.shape[0], 3), np.complex128)
        H = np.zeros((r.shape[0], 3), np.complex128)
        nat.check(nat.lib().spk_quad_scattered_field(
            self._h, r.shape[0], _dp(r), _dp(x.view(np.float64)), _dp(E.view(np.float64)),
            _dp(H.view(np.float64))))
        return E, H


def topology_probe(Lx, Ly, nx, ny, z):
    """Host-only mesh and edge numbering of spk_quad_setup (no GPU needed):
    (pp, vert8 1-based, edge_p5)."""
    z = np.ascontiguousarray(z, dtype=np.float64).reshape(-1)
    L = nat.lib()
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    nat.check(L.spk_quad_topology_probe(Lx, Ly, nx, ny, _dp(z), C.byref(a), C.byref(b),
                                        C.byref(c), None, None, None))
    pp = np.zeros((c.value, 3))
    v = np.zeros((b.value, 8), np.int32)
    e = np.zeros((a.value, 5), np.int32)
    nat.check(L.spk_quad_topology_probe(Lx, Ly, nx, ny, _dp(z), C.byref(a), C.byref(b),
                                        C.byref(c), _dp(pp), _ip(v), _ip(e)))
    return pp, v, e
