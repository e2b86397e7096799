"""ctypes binding of the CPU restatement in ``spk_oracle.c``.

Test infrastructure only: the checker the GPU product is compared against.
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libspk_oracle.so")

FORMULATIONS = {"PMCHWT": 1, "MCTF": 2, "ICTF": 3, "MCTF2": 4, "PMCHWT-S": 5,
                "JMCFIE": 6, "CNF": 7, "MNMF": 8}


class _Params(C.Structure):
    _fields_ = [("lambda_", C.c_double), ("eps_r1_re", C.c_double), ("eps_r1_im", C.c_double),
                ("eps_r2_re", C.c_double), ("eps_r2_im", C.c_double), ("formulation", C.c_int),
                ("normal_convention", C.c_int), ("ngp", C.c_int), ("s_opt", C.c_int),
                ("K", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        L.spo_create.restype = C.c_void_p
        L.spo_create.argtypes = [C.POINTER(_Params), C.c_int, dp, C.c_int, ip]
        L.spo_destroy.argtypes = [C.c_void_p]
        L.spo_n_edges.argtypes = [C.c_void_p]
        L.spo_get_pairs.argtypes = [C.c_void_p, ip, ip]
        L.spo_get_fn.argtypes = [C.c_void_p, C.c_int, C.c_int, dp]
        L.spo_near_block.argtypes = [C.c_void_p, C.c_int, C.c_int, dp]
        L.spo_edge_adjacent_kt.argtypes = [C.c_void_p] + [C.c_int] * 5 + [dp]
        L.spo_tree_build.argtypes = [C.c_void_p]
        L.spo_tree_info.argtypes = [C.c_void_p, ip, ip, dp, dp]
        L.spo_edge_leaf.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
        L.spo_e2_of_edge.argtypes = [C.c_void_p, C.c_int, ip, C.c_int]
        L.spo_i4.argtypes = [C.c_void_p, C.c_int, C.c_int64, C.POINTER(C.c_int64), C.c_int]
        L.spo_near_csr_build.restype = C.c_int64
        L.spo_near_csr_build.argtypes = [C.c_void_p]
        L.spo_near_csr_get.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.spo_near_fill.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.spo_near_values.argtypes = [C.c_void_p, dp]
        L.spo_mlfmm_setup.argtypes = [C.c_void_p, C.c_int]
        L.spo_matvec.argtypes = [C.c_void_p, dp, dp, C.c_int]
        L.spo_lagrange_interp.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp]
        L.spo_set_level_K.argtypes = [C.c_void_p, C.c_int, ip, C.c_int]
        L.spo_mlfmm_setup_levels.argtypes = [C.c_void_p, C.c_int]
        L.spo_rhs_plane.argtypes = [C.c_void_p, dp, dp, dp]
        L.spo_scattered_field.argtypes = [C.c_void_p, C.c_int, dp, dp, dp, dp]
        L.spo_gmres.argtypes = [C.c_void_p, dp, dp, C.c_int, C.c_int, C.c_double, C.c_int,
                                C.c_int, dp, ip, dp]
        L.spq_create.restype = C.c_void_p
        L.spq_create.argtypes = [C.c_double] * 5 + [C.c_double, C.c_double, C.c_int, C.c_int, dp]
        L.spq_destroy.argtypes = [C.c_void_p]
        for f in ("spq_n_edges", "spq_n_elements", "spq_n_nodes"):
            getattr(L, f).argtypes = [C.c_void_p]
        L.spq_nodes.argtypes = [C.c_void_p, dp]
        L.spq_elements.argtypes = [C.c_void_p, ip]
        L.spq_edges.argtypes = [C.c_void_p, ip]
        L.spq_edge_element.argtypes = [C.c_void_p, C.c_int, C.c_int, dp]
        L.spq_self_point.argtypes = [C.c_void_p, dp, C.c_int, C.c_double, C.c_double, C.c_int,
                                     dp]
        L.spq_assemble.argtypes = [C.c_void_p, C.c_int, dp, C.c_int]
        L.spq_geometry_probe.argtypes = [C.c_double, C.c_double, dp, dp]
        L.spq_regular_pair.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, dp]
        L.spq_rhs_plane.argtypes = [C.c_void_p, C.c_int, dp, dp, dp]
        L.spq_scattered_field.argtypes = [C.c_void_p, C.c_int, dp, dp, C.c_int, dp, dp]
        L.spo_gmres_prec.argtypes = [C.c_void_p, dp, dp, C.c_int, C.c_int, C.c_double, C.c_int,
                                     C.c_int, C.c_int, dp, ip, dp, dp]
        L.spo_gmres_dense.argtypes = [C.c_int, dp, dp, dp, C.c_int, C.c_int, C.c_double, C.c_int,
                                      dp, ip, dp]
        L.spo_gmres_dense_prec.argtypes = [C.c_int, dp, dp, dp, dp, C.c_int, C.c_int, C.c_double,
                                           C.c_int, dp, ip, dp, dp]
        L.spo_quadrature_tri.argtypes = [C.c_int, dp, dp, dp]
        L.spo_iq_l.restype = C.c_double
        L.spo_iq_l.argtypes = [C.c_int, dp, dp, dp]
        L.spo_iq_s_all.argtypes = [dp, dp, dp, dp, C.c_int, dp]
        for f in (L.spo_k2q, L.spo_k3q, L.spo_k4q):
            f.argtypes = [C.c_int, dp, dp, C.c_int, dp]
        L.spo_gauss_legendre.argtypes = [C.c_int, dp, dp]
        L.spo_hankel2_sph.argtypes = [C.c_int, C.c_double, C.c_double, dp]
        L.spo_legendre_p.argtypes = [C.c_int, C.c_double, dp]
        L.spo_tl.argtypes = [C.c_int, C.c_double, C.c_double, dp, dp]
        L.spo_dense_fill.argtypes = [C.c_void_p, dp, C.c_int]
        L.spo_get_pattern.argtypes = [C.c_void_p, C.c_int, dp]
        L.spo_tree_uindex.restype = C.c_uint64
        L.spo_tree_uindex.argtypes = [dp, C.c_int]
        L.spo_tree_neighbours.argtypes = [C.c_int64, C.c_int, C.POINTER(C.c_int64)]
        L.spo_tree_centre.argtypes = [C.c_int64, C.c_int, dp]
        L.spo_tree_level_max.argtypes = [C.c_int, dp, C.c_int, C.c_int]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class Oracle:
    """One mesh + material + formulation, as the reference's module state."""

    def __init__(self, pp, tt_1based, lam, eps_r2, eps_r1=1.0 + 0j, formulation="PMCHWT",
                 normal_convention=1, ngp=3, s_opt=250, K=19):
        L = lib()
        self.pp = np.ascontiguousarray(pp, dtype=np.float64).reshape(-1, 3)
        self.tt = np.ascontiguousarray(tt_1based, dtype=np.int32).reshape(-1, 3)
        f = FORMULATIONS[formulation] if isinstance(formulation, str) else int(formulation)
        self.params = _Params(lam, complex(eps_r1).real, complex(eps_r1).imag,
                              complex(eps_r2).real, complex(eps_r2).imag, f,
                              normal_convention, ngp, s_opt, K)
        self.h = L.spo_create(C.byref(self.params), self.pp.shape[0], _d(self.pp),
                              self.tt.shape[0], _i(self.tt))
        self.N = L.spo_n_edges(self.h)
        self._csr = None

    def close(self):
        if self.h:
            lib().spo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def pairs(self):
        c = np.zeros((self.N, 2), np.int32)
        e = np.zeros((self.N, 2), np.int32)
        lib().spo_get_pairs(self.h, _i(c), _i(e))
        return c, e

    def fn(self, m, ot):
        o = np.zeros(19)
        lib().spo_get_fn(self.h, m, ot, _d(o))
        return o

    def near_block(self, m, n):
        o = np.zeros(8)
        lib().spo_near_block(self.h, m, n, _d(o))
        return (o[0::2] + 1j * o[1::2]).reshape(2, 2)

    def edge_adjacent_kt(self, m, t, n, s, md):
        """The 12x12-rule K of side t of edge m against side s of edge n in
        medium md (1 outside, 2 inside): (K, its double integral, x_nt)."""
        o = np.zeros(5)
        lib().spo_edge_adjacent_kt(self.h, m, t, n, s, md, _d(o))
        return o[0] + 1j * o[1], o[2] + 1j * o[3], o[4]

    def tree(self):
        L = lib()
        lmax = L.spo_tree_build(self.h)
        a = C.c_int()
        b = C.c_int()
        D = np.zeros(1)
        xm = np.zeros(3)
        L.spo_tree_info(self.h, C.byref(a), C.byref(b), _d(D), _d(xm))
        leaf = np.zeros(self.N, np.int64)
        L.spo_edge_leaf(self.h, leaf.ctypes.data_as(C.POINTER(C.c_int64)))
        return dict(l_min=a.value, l_max=b.value, D=D[0], xmin=xm, leaf=leaf)

    def e2(self, m):
        L = lib()
        k = L.spo_e2_of_edge(self.h, m, None, 0)
        o = np.zeros(k, np.int32)
        L.spo_e2_of_edge(self.h, m, _i(o), k)
        return o

    def i4(self, level, n):
        o = np.zeros(216, np.int64)
        k = lib().spo_i4(self.h, level, n, o.ctypes.data_as(C.POINTER(C.c_int64)), 216)
        return o[:k]

    def near_csr(self, fill=True, nthreads=0, rows=None):
        L = lib()
        nnz = L.spo_near_csr_build(self.h)
        ia = np.zeros(self.N + 1, np.int32)
        ja = np.zeros(nnz, np.int32)
        L.spo_near_csr_get(self.h, ia.ctypes.data_as(C.POINTER(C.c_int32)),
                           ja.ctypes.data_as(C.POINTER(C.c_int32)))
        vals = None
        if fill:
            lo, hi = (0, self.N) if rows is None else rows
            L.spo_near_fill(self.h, lo, hi, nthreads)
            v = np.zeros(nnz * 8)
            L.spo_near_values(self.h, _d(v))
            vals = (v[0::2] + 1j * v[1::2]).reshape(nnz, 4)
        self._csr = (ia, ja, vals)
        return ia, ja, vals

    def near_fill_rows(self, lo, hi, nthreads=0):
        lib().spo_near_fill(self.h, lo, hi, nthreads)

    def dense(self, nthreads=0):
        Z = np.zeros((2 * self.N, 2 * self.N), np.complex128)
        lib().spo_dense_fill(self.h, _d(Z.view(np.float64)), nthreads)
        return Z

    def mlfmm_setup(self, nthreads=0, level_K=None, rr="adjoint"):
        """Patterns and TL tables; level_K: truncation number per level
        (index = level; 0 or missing = the default K), non-increasing from
        the coarsest level to the leaves; rr: "reference" (RR interpolates
        the local coefficients, as the reference's code would) or "adjoint"
        (RR applies the transpose of SS's interpolation)."""
        if level_K is not None:
            k = np.ascontiguousarray(level_K, dtype=np.int32)
            lib().spo_set_level_K(self.h, k.size, _i(k), {"reference": 0, "adjoint": 1}[rr])
        if lib().spo_mlfmm_setup_levels(self.h, nthreads) != 0:
            raise ValueError("a level's truncation number exceeds its parent's")
        self.level_K = None if level_K is None else list(level_K)

    def pattern(self, m):
        K = self.params.K
        if getattr(self, "level_K", None) is not None:
            lmax = self.tree()["l_max"]
            if lmax < len(self.level_K) and self.level_K[lmax] > 0:
                K = self.level_K[lmax]
        S = 2 * K * K + 2
        o = np.zeros(8 * S)
        lib().spo_get_pattern(self.h, m, _d(o))
        return (o[0::2] + 1j * o[1::2]).reshape(4, S)

    def matvec(self, x, nthreads=0):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros(2 * self.N, np.complex128)
        lib().spo_matvec(self.h, _d(x.view(np.float64)), _d(y.view(np.float64)), nthreads)
        return y

    def rhs_plane(self, k_in, e_in):
        k = np.ascontiguousarray(k_in, dtype=np.float64)
        e = np.ascontiguousarray(e_in, dtype=np.float64)
        b = np.zeros(2 * self.N, np.complex128)
        lib().spo_rhs_plane(self.h, _d(k), _d(e), _d(b.view(np.float64)))
        return b

    def scattered_field(self, points, x):
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        x = np.ascontiguousarray(x, dtype=np.complex128)
        E = np.zeros((pts.shape[0], 3), np.complex128)
        H = np.zeros((pts.shape[0], 3), np.complex128)
        lib().spo_scattered_field(self.h, pts.shape[0], _d(pts), _d(x.view(np.float64)),
                                  _d(E.view(np.float64)), _d(H.view(np.float64)))
        return E, H

    def gmres_left_jacobi(self, b, maxit=650, restart=0, tol=1e-3, orth=1, nthreads=0):
        b = np.ascontiguousarray(b, dtype=np.complex128)
        x = np.zeros(2 * self.N, np.complex128)
        hist = np.zeros(maxit)
        it = C.c_int()
        be = np.zeros(1)
        tbe = np.zeros(1)
        info = lib().spo_gmres_prec(self.h, _d(b.view(np.float64)), _d(x.view(np.float64)), maxit,
                                    restart, tol, orth, 1, nthreads, _d(hist), C.byref(it), _d(be),
                                    _d(tbe))
        return x, dict(info=info, iters=it.value, be=be[0], tbe=tbe[0], hist=hist[:it.value])

    def gmres(self, b, maxit=650, restart=0, tol=1e-3, orth=1, nthreads=0, x0=None):
        b = np.ascontiguousarray(b, dtype=np.complex128)
        x = np.zeros(2 * self.N, np.complex128) if x0 is None else np.array(x0, np.complex128)
        hist = np.zeros(maxit)
        it = C.c_int()
        be = np.zeros(1)
        info = lib().spo_gmres(self.h, _d(b.view(np.float64)), _d(x.view(np.float64)), maxit,
                               restart, tol, orth, nthreads, _d(hist), C.byref(it), _d(be))
        return x, dict(info=info, iters=it.value, be=be[0], hist=hist[:it.value])


def lagrange_interp(values, Km, Kn, p=3):
    """Lagrange_Interpolation_New (src/lib_sie_mlfmm.f90:268-390) for one
    current: values (4, 2 Km^2 + 2) complex, channels (theta k1, theta k2,
    phi k1, phi k2), the last two samples the Cartesian x / y components at
    theta = 0 and pi; returns (4, 2 Kn^2 + 2)."""
    v = np.ascontiguousarray(values, dtype=np.complex128).reshape(4, 2 * Km * Km + 2)
    out = np.zeros((4, 2 * Kn * Kn + 2), np.complex128)
    lib().spo_lagrange_interp(int(p), int(Km), int(Kn), _d(v.view(np.float64)),
                              _d(out.view(np.float64)))
    return out


def sample_grid(K):
    """Khat_and_TRF_ExtendedK (src/lib_sie_mlfmm.f90:190-249): (khat, theta
    unit vectors, phi unit vectors, weights) of the 2K^2 + 2 samples."""
    x, w = np.polynomial.legendre.leggauss(K)
    S = 2 * K * K + 2
    kh, tt, tp, trf = np.zeros((S, 3)), np.zeros((S, 3)), np.zeros((S, 3)), np.zeros(S)
    dphi = np.pi / K
    for j in range(K):
        st = np.sin(np.arccos(x[j]))
        for k in range(2 * K):
            ph, i = k * dphi, k + j * 2 * K
            kh[i] = [st * np.cos(ph), st * np.sin(ph), x[j]]
            tt[i] = [x[j] * np.cos(ph), x[j] * np.sin(ph), -st]
            tp[i] = [-np.sin(ph), np.cos(ph), 0.0]
            trf[i] = w[j] * dphi
    kh[S - 2], kh[S - 1] = [0, 0, 1], [0, 0, -1]
    tt[S - 2] = tt[S - 1] = [1, 0, 0]
    tp[S - 2] = tp[S - 1] = [0, 1, 0]
    return kh, tt, tp, trf


def gmres_dense(A, b, maxit, restart=0, tol=1e-6, orth=1):
    A = np.asfortranarray(A, dtype=np.complex128)
    b = np.ascontiguousarray(b, dtype=np.complex128)
    n = b.shape[0]
    x = np.zeros(n, np.complex128)
    hist = np.zeros(maxit)
    it = C.c_int()
    be = np.zeros(1)
    Af = np.ascontiguousarray(A.T.reshape(-1))  # column-major flattening
    info = lib().spo_gmres_dense(n, _d(Af.view(np.float64)), _d(b.view(np.float64)),
                                 _d(x.view(np.float64)), maxit, restart, tol, orth, _d(hist),
                                 C.byref(it), _d(be))
    return x, dict(info=info, iters=it.value, be=be[0], hist=hist[:it.value])


def gmres_dense_left_jacobi(A, diag, b, maxit, restart=0, tol=1e-6, orth=1):
    A = np.asfortranarray(A, dtype=np.complex128)
    b = np.ascontiguousarray(b, dtype=np.complex128)
    dg = np.ascontiguousarray(diag, dtype=np.complex128)
    n = b.shape[0]
    x = np.zeros(n, np.complex128)
    hist = np.zeros(maxit)
    it = C.c_int()
    be = np.zeros(1)
    tbe = np.zeros(1)
    Af = np.ascontiguousarray(A.T.reshape(-1))
    info = lib().spo_gmres_dense_prec(n, _d(Af.view(np.float64)), _d(dg.view(np.float64)),
                                      _d(b.view(np.float64)), _d(x.view(np.float64)), maxit,
                                      restart, tol, orth, _d(hist), C.byref(it), _d(be), _d(tbe))
    return x, dict(info=info, iters=it.value, be=be[0], tbe=tbe[0], hist=hist[:it.value])


class QuadOracle:
    """Restatement of the reference's 8-node quadrilateral fill (spk_quad.c)
    on the height-field mesh of sorting_quad."""

    def __init__(self, lam, eps_r2, Lx, Ly, nx, ny, z, eps_r1=1.0):
        L = lib()
        e1, e2 = complex(eps_r1), complex(eps_r2)
        z = np.ascontiguousarray(z, dtype=np.float64).reshape(-1)
        assert z.size == (2 * nx + 1) * (2 * ny + 1)
        self._z = z
        self.h = L.spq_create(lam, e1.real, e1.imag, e2.real, e2.imag, Lx, Ly, nx, ny, _d(z))
        if not self.h:
            raise ValueError("bad quad mesh")
        self.M = L.spq_n_edges(self.h)
        self.n_el = L.spq_n_elements(self.h)
        self.n_nodes = L.spq_n_nodes(self.h)

    def __del__(self):
        try:
            lib().spq_destroy(self.h)
        except Exception:
            pass

    def nodes(self):
        pp = np.zeros((self.n_nodes, 3))
        lib().spq_nodes(self.h, _d(pp))
        return pp

    def elements(self):
        v = np.zeros((self.n_el, 8), np.int32)
        lib().spq_elements(self.h, v.ctypes.data_as(C.POINTER(C.c_int)))
        return v

    def edges(self):
        e = np.zeros((self.M, 5), np.int32)
        lib().spq_edges(self.h, e.ctypes.data_as(C.POINTER(C.c_int)))
        return e

    def edge_element(self, s, t):
        """Sum_aa(3, 10) of impedance_edge_versus_element, 1-based s, t."""
        o = np.zeros(60)
        lib().spq_edge_element(self.h, int(s), int(t), _d(o))
        return (o[0::2] + 1j * o[1::2]).reshape(3, 10)

    def element_nodes(self, e):
        """The 8 nodes (8, 3) of element e (0-based), in the element's order."""
        return self.nodes()[self.elements()[e] - 1]

    def self_point(self, El, pI, mu, nu, ngp=15):
        """Integration_KABCD_Taylor_ST_4FEdge at (mu, nu) of element El (8, 3)
        for test function pI (1-based): (raw (5, 10): KA, KBE, KBH with their
        log terms, KCE, KCH; combined (3, 10))."""
        El = np.ascontiguousarray(El, dtype=np.float64).reshape(24)
        o = np.zeros(160)
        lib().spq_self_point(self.h, _d(El), int(pI), float(mu), float(nu), int(ngp), _d(o))
        z = o[0::2] + 1j * o[1::2]
        return z[:50].reshape(5, 10), z[50:].reshape(3, 10)

    def regular_pair(self, s, t, ngp):
        o = np.zeros(60)
        lib().spq_regular_pair(self.h, int(s), int(t), int(ngp), _d(o))
        return (o[0::2] + 1j * o[1::2]).reshape(3, 10)

    def assemble(self, formulation="PMCHWT", nthreads=0):
        f = {"PMCHWT": 1, "MCTF": 2, "ICTF": 3}[formulation]
        MM = np.zeros((2 * self.M, 2 * self.M), np.complex128)
        lib().spq_assemble(self.h, f, _d(MM.view(np.float64)), nthreads)
        return MM

    def rhs_plane(self, k_in, e_in, formulation="PMCHWT"):
        """get_vector_b_quad with a plane wave: b[2M] complex."""
        f = {"PMCHWT": 1, "MCTF": 2, "ICTF": 3}[formulation]
        k = np.ascontiguousarray(k_in, dtype=np.float64)
        e = np.ascontiguousarray(e_in, dtype=np.float64)
        b = np.zeros(2 * self.M, np.complex128)
        lib().spq_rhs_plane(self.h, f, _d(k), _d(e), _d(b.view(np.float64)))
        return b


    def scattered_field(self, points, x, ngp=5):
        """E, H (npts, 3) of a solution x[2M] (Integration_scattered_field_quad
        over all elements on the solution's element-local coefficients)."""
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        xx = np.ascontiguousarray(x, dtype=np.complex128)
        E = np.zeros((pts.shape[0], 3), np.complex128)
        H = np.zeros((pts.shape[0], 3), np.complex128)
        lib().spq_scattered_field(self.h, pts.shape[0], _d(pts), _d(xx.view(np.float64)), int(ngp),
                                  _d(E.view(np.float64)), _d(H.view(np.float64)))
        return E, H


def quad_geometry(k, e, El):
    """Element geometry probe: dict of shape functions, derivatives, JJ, nq,
    the 10 edge functions and their n.curl at (k, e)."""
    El = np.ascontiguousarray(El, dtype=np.float64).reshape(24)
    o = np.zeros(68)
    lib().spq_geometry_probe(float(k), float(e), _d(El), _d(o))
    return dict(N=o[0:8], dk=o[8:16], de=o[16:24], JJ=o[24], nq=o[25:28],
                fn=o[28:58].reshape(10, 3), nrot=o[58:68])


def quadrature_tri(ngp):
    a = np.zeros(20)
    b = np.zeros(20)
    w = np.zeros(20)
    lib().spo_quadrature_tri(ngp, _d(a), _d(b), _d(w))
    return a[:ngp], b[:ngp], w[:ngp]


def iq_l(q, ra, p1, p2):
    return lib().spo_iq_l(q, _d(np.asarray(ra, float)), _d(np.asarray(p1, float)),
                          _d(np.asarray(p2, float)))


def iq_s_all(ra, corners, nq, mid, iqs_l=False):
    o = np.zeros(3)
    lib().spo_iq_s_all(_d(np.asarray(ra, float)), _d(np.ascontiguousarray(corners, float)),
                       _d(np.asarray(nq, float)), _d(np.asarray(mid, float)), int(iqs_l), _d(o))
    return o


def kq(which, q, ra, fn19, iqs_l=False):
    o = np.zeros(3)
    f = {2: lib().spo_k2q, 3: lib().spo_k3q, 4: lib().spo_k4q}[which]
    f(q, _d(np.asarray(ra, float)), _d(np.ascontiguousarray(fn19, float)), int(iqs_l), _d(o))
    return o


def gauss_legendre(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().spo_gauss_legendre(n, _d(x), _d(w))
    return x, w


def hankel2_sph(n, z):
    o = np.zeros(2 * n)
    lib().spo_hankel2_sph(n, complex(z).real, complex(z).imag, _d(o))
    return o[0::2] + 1j * o[1::2]


def legendre_p(n, x):
    o = np.zeros(n)
    lib().spo_legendre_p(n, x, _d(o))
    return o


def tl(K, k, r):
    S = 2 * K * K + 2
    o = np.zeros(2 * S)
    lib().spo_tl(K, complex(k).real, complex(k).imag, _d(np.asarray(r, float)), _d(o))
    return o[0::2] + 1j * o[1::2]


def tree_uindex(xs, l):
    return int(lib().spo_tree_uindex(_d(np.asarray(xs, float)), l))


def tree_neighbours(n, l):
    o = np.zeros(27, np.int64)
    k = lib().spo_tree_neighbours(n, l, o.ctypes.data_as(C.POINTER(C.c_int64)))
    return o[:k]


def tree_centre(n, l):
    o = np.zeros(3)
    lib().spo_tree_centre(n, l, _d(o))
    return o


def tree_level_max(xs, s, l_th):
    xs = np.ascontiguousarray(xs, dtype=np.float64).reshape(-1, 3)
    return lib().spo_tree_level_max(xs.shape[0], _d(xs), s, l_th)
