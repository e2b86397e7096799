"""Host-side mirror of SpeckleSim's triangle MLFMM-GMRES module API.

``SieTriMlfmm`` follows ``lib_sie_tri_mlfmm_interface`` and
``lib_sie_solver_gmres_interface`` of the reference
(src/lib_sie_tri_mlfmm_interface.f90:18-26, src/lib_sie_solver_gmres_interface.f90:27-29):

=====================================  ==========================================
reference routine                      method here (GPU through libspeckle_gpu.so)
=====================================  ==========================================
lib_sie_ml_fmm_precalculation_tri      precalculation_tri(pp, tt)
lib_sie_ml_fmm_calculate_vector_b_tri  calculate_vector_b_tri(vector_x)
get_vector_b_tri (plane wave)          get_vector_b_tri(k_in, e_in)
lib_sie_ml_fmm_get_init_vector_x_tri   get_init_vector_x_tri(x_initial)
lib_math_solver_gmres_run              gmres_run(vector_b, ...)
lib_sie_tri_get_impedance_edges        impedance_edges(m, n)   (0-based)
=====================================  ==========================================

Vectors are complex128 numpy arrays of length 2N in the reference's order
[xe(1:N); xh(1:N)].  Mesh triangle indices are 1-based (``tt_*`` files).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as nat

FORMULATIONS = {"PMCHWT": 1, "MCTF": 2, "ICTF": 3, "MCTF2": 4, "PMCHWT-S": 5,
                "JMCFIE": 6, "CNF": 7, "MNMF": 8}
ORTHOGONALIZATION = {"MGS": 0, "IMGS": 1, "CGS": 2, "ICGS": 3}


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


PRECOND_NEAR_FIELD = 2
PRECOND_BLOCK_LU = 3


def _gmres_params(max_iterations, restart, tolerance, orthogonalization, use_initial_guess,
                  refinement_steps, refinement_tolerance, precondition_left, near_field,
                  inner_iterations):
    if precondition_left and near_field:
        raise ValueError("precondition_left and near_field exclude each other")
    # near_field: True for the inner near-field GMRES, "block_lu" for the
    # block-Jacobi inverse of the leaf groups' near-field matrices
    if near_field == "block_lu":
        prec = PRECOND_BLOCK_LU
    elif near_field:
        prec = PRECOND_NEAR_FIELD
    else:
        prec = int(bool(precondition_left))
    return nat.GmresParams(max_iterations, restart, tolerance, ORTHOGONALIZATION[orthogonalization],
                           int(use_initial_guess), refinement_steps, refinement_tolerance, prec,
                           inner_iterations)


def plane_wave(theta_in=0.0, phi_in=0.0, k_direction=1, polarization="TM"):
    """Incident plane wave of set_illumination_parameter (main.f90:245-297):
    k_in = (sin t cos p, sin t sin p, sign cos t) with sign = +1 for
    calc_p(7)=1 and -1 for calc_p(7)=2, E_in from field_rotation_matrix
    (main.f90:301-324) with TM -> (1,0), TE -> (0,1)."""
    sign = 1.0 if k_direction == 1 else -1.0
    k = np.array([np.sin(theta_in) * np.cos(phi_in), np.sin(theta_in) * np.sin(phi_in),
                  sign * np.cos(theta_in)])
    ein = (1.0, 0.0) if polarization.upper() == "TM" else (0.0, 1.0)
    a, b, c, d = np.sin(theta_in), np.cos(theta_in), np.sin(phi_in), np.cos(phi_in)
    rm = np.array([[d * d * b + c * c, c * d * (b - 1)], [c * d * (b - 1), c * c * b + d * d],
                   [a * d, a * c]])
    return k, rm @ np.asarray(ein)


class SieTriMlfmm:
    """One mesh on one GPU: near-field blocks, MLFMM and GMRES in HBM."""

    def __init__(self, wavelength, eps_r2, eps_r1=1.0 + 0.0j, formulation="PMCHWT",
                 normal_convention=1, ng=3, tree_s_opt=250, truncation_number=19,
                 precision="fp64", device=0):
        L = nat.lib()
        nat.check_single_hip_runtime()
        f = FORMULATIONS[formulation] if isinstance(formulation, str) else int(formulation)
        prec = {"fp64": 0, "mixed": 1}[precision]
        eps_r1, eps_r2 = complex(eps_r1), complex(eps_r2)
        self.params = nat.Params(wavelength, eps_r1.real, eps_r1.imag, eps_r2.real, eps_r2.imag,
                                 f, normal_convention, ng, tree_s_opt, truncation_number, prec,
                                 device)
        h = C.c_void_p()
        nat.check(L.spk_create(C.byref(self.params), C.byref(h)))
        self._h = h
        self.N = 0

    def close(self):
        if getattr(self, "_h", None):
            nat.lib().spk_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- lib_sie_ml_fmm_precalculation_tri ---------------------------------
    def precalculation_tri(self, pp, tt):
        pp = np.ascontiguousarray(pp, dtype=np.float64).reshape(-1, 3)
        tt = np.ascontiguousarray(tt, dtype=np.int32).reshape(-1, 3)
        nat.check(nat.lib().spk_setup_tri(self._h, pp.shape[0], _dp(pp), tt.shape[0],
                                          tt.ctypes.data_as(C.POINTER(C.c_int))))
        n = C.c_int64()
        nat.check(nat.lib().spk_num_edges(self._h, C.byref(n)))
        self.N = n.value
        return self

    def precalculation_quad(self, Lx, Ly, nx, ny, z):
        """The MLFMM operator on an 8-node quadrilateral height-field surface
        (spk_setup_quad): Lx x Ly, nx x ny elements, z[(2nx+1)(2ny+1)] node
        heights as the reference's height file orders them (x index fastest;
        sorting_quad, src/lib_sie_quad_data_container.f90:321-403).  N is
        then the number of edge functions M; the products, right-hand sides
        (get_vector_b_quad), solves and scattered fields of this object work
        on the quadrilateral system."""
        z = np.ascontiguousarray(z, dtype=np.float64).reshape(-1)
        if z.size != (2 * nx + 1) * (2 * ny + 1):
            raise ValueError("z must hold (2nx+1)(2ny+1) heights")
        nat.check(nat.lib().spk_setup_quad(self._h, float(Lx), float(Ly), int(nx), int(ny), _dp(z)))
        n = C.c_int64()
        nat.check(nat.lib().spk_num_edges(self._h, C.byref(n)))
        self.N = n.value
        return self

    # -- lib_sie_ml_fmm_calculate_vector_b_tri ------------------------------
    def calculate_vector_b_tri(self, vector_x):
        x = np.ascontiguousarray(vector_x, dtype=np.complex128)
        if x.shape != (2 * self.N,):
            raise ValueError(f"vector_x must have length 2N = {2 * self.N}")
        y = np.empty_like(x)
        nat.check(nat.lib().spk_matvec(self._h, _dp(x.view(np.float64)), _dp(y.view(np.float64))))
        return y

    def calculate_vector_b_tri_c64(self, vector_x):
        """The complex64 operator of a precision="mixed" handle."""
        x = np.ascontiguousarray(vector_x, dtype=np.complex64)
        if x.shape != (2 * self.N,):
            raise ValueError(f"vector_x must have length 2N = {2 * self.N}")
        y = np.empty_like(x)
        fp = C.POINTER(C.c_float)
        nat.check(nat.lib().spk_matvec_c64(self._h, x.view(np.float32).ctypes.data_as(fp),
                                           y.view(np.float32).ctypes.data_as(fp)))
        return y

    def matvec_parts(self, vector_x):
        x = np.ascontiguousarray(vector_x, dtype=np.complex128)
        yn = np.empty_like(x)
        yf = np.empty_like(x)
        nat.check(nat.lib().spk_matvec_parts(self._h, _dp(x.view(np.float64)),
                                             _dp(yn.view(np.float64)), _dp(yf.view(np.float64))))
        return yn, yf

    def calculate_vector_b_tri_device(self, d_x, d_y):
        """Device-pointer matvec (e.g. torch complex128 tensors' data_ptr())."""
        nat.check(nat.lib().spk_matvec_device(self._h, C.c_void_p(int(d_x)), C.c_void_p(int(d_y))))

    def calculate_vector_b_tri_batch_device(self, d_xs, d_ys, c64=False):
        """Batched device matvec: d_xs, d_ys are sequences of R <= 16 device
        pointers (2N complex128 vectors, or complex64 with c64=True on a
        precision="mixed" handle)."""
        R = len(d_xs)
        if len(d_ys) != R:
            raise ValueError("d_xs and d_ys must have the same length")
        xs = (C.c_void_p * R)(*[int(p) for p in d_xs])
        ys = (C.c_void_p * R)(*[int(p) for p in d_ys])
        f = nat.lib().spk_matvec_batch_c64_device if c64 else nat.lib().spk_matvec_batch_device
        nat.check(f(self._h, R, xs, ys))

    def calculate_vector_b_tri_batch(self, X, c64=False):
        """Host convenience: X is (R, 2N); returns Z X row by row through the
        batched device matvec."""
        from ._native import DeviceBuffer

        dt = np.complex64 if c64 else np.complex128
        X = np.ascontiguousarray(X, dtype=dt)
        R = X.shape[0]
        nb = 2 * self.N * X.itemsize
        xb = [DeviceBuffer(nb) for _ in range(R)]
        yb = [DeviceBuffer(nb) for _ in range(R)]
        for r in range(R):
            xb[r].upload(X[r])
        self.calculate_vector_b_tri_batch_device([b.ptr for b in xb], [b.ptr for b in yb], c64)
        return np.stack([yb[r].download(dt) for r in range(R)])

    # -- get_vector_b_tri / get_init_vector_x_tri --------------------------
    def get_vector_b_tri(self, k_in, e_in):
        k = np.ascontiguousarray(k_in, dtype=np.float64)
        e = np.ascontiguousarray(e_in, dtype=np.float64)
        b = np.empty(2 * self.N, np.complex128)
        nat.check(nat.lib().spk_rhs_plane_wave(self._h, _dp(k), _dp(e), _dp(b.view(np.float64))))
        return b

    def get_vector_b_tri_device(self, k_in, e_in, d_b):
        """Plane-wave right-hand side into a device buffer of 2N complex128."""
        k = np.ascontiguousarray(k_in, dtype=np.float64)
        e = np.ascontiguousarray(e_in, dtype=np.float64)
        nat.check(nat.lib().spk_rhs_plane_wave_device(self._h, _dp(k), _dp(e), C.c_void_p(int(d_b))))

    def get_init_vector_x_tri(self, x_initial=0.0):
        return np.full(2 * self.N, complex(x_initial), dtype=np.complex128)

    # -- the transposed operator (the adjoint system) ------------------------
    def set_transposed(self, transposed=True):
        """Every later matvec and solve uses Z^T (or Z again)."""
        nat.check(nat.lib().spk_set_transposed(self._h, int(bool(transposed))))
        return self

    @property
    def transposed(self):
        t = C.c_int()
        nat.check(nat.lib().spk_get_transposed(self._h, C.byref(t)))
        return bool(t.value)

    # -- per-level truncation numbers ----------------------------------------
    def set_level_truncation(self, K_by_level, rr="anterpolate"):
        """Truncation number per octree level (index = level, 0 = the
        handle's default) for the next precalculation_tri; rr: "reference"
        (RR interpolates the parent's local coefficients, as the reference's
        code does) or "anterpolate" (RR applies the transpose of SS's
        interpolation to folded coefficients).  FP64 handles."""
        k = np.ascontiguousarray(K_by_level, dtype=np.int32).reshape(-1)
        mode = {"reference": 0, "anterpolate": 1}[rr]
        nat.check(nat.lib().spk_set_level_truncation(self._h, k.size,
                                                     k.ctypes.data_as(C.POINTER(C.c_int)), mode))
        return self

    # -- run-time tuning --------------------------------------------------------
    def autotune(self, angles, tolerance=1e-6, max_iterations=650):
        """Tunes the near-field preconditioned solver on the current mesh
        (spk_autotune): the batched fp8 near product's row tiles per
        workgroup and the inner solve length, timed on a batched solve of
        ``angles`` (up to 16 (k_in, e_in) pairs).  The choice stays in the
        handle; returns it with the measured times."""
        kin = np.ascontiguousarray([a[0] for a in angles], dtype=np.float64).reshape(-1, 3)
        ein = np.ascontiguousarray([a[1] for a in angles], dtype=np.float64).reshape(-1, 3)
        gp = _gmres_params(max_iterations, 0, tolerance, "ICGS", False, 0, 0.0, False, True, 0)
        t = nat.Tuning()
        nat.check(nat.lib().spk_autotune(self._h, kin.shape[0], _dp(kin), _dp(ein), C.byref(gp),
                                         C.byref(t)))
        return dict(tiles_per_group=t.tiles_per_group, inner_iterations=t.inner_iterations,
                    tile_us=list(t.tile_us),
                    inner_ms={t.inner_tried[i]: t.inner_ms[i] for i in range(t.n_inner)})

    def set_tuning(self, tiles_per_group=0, inner_iterations=0):
        """Applies tuned values (0: the default)."""
        nat.check(nat.lib().spk_set_tuning(self._h, int(tiles_per_group), int(inner_iterations)))
        return self

    # -- adjoint gradient with respect to node heights ------------------------
    def adjoint_height_gradient(self, k_in, e_in, points, weights, nodes, dz=1e-6,
                                tolerance=1e-10, max_iterations=650, near_field=True,
                                inner_iterations=0, return_terms=False):
        """J = sum_p weights[p] |E(points[p])|^2 of the plane-wave solution
        and dJ/dz of the mesh nodes ``nodes`` (0-based), by one forward and
        one transposed solve and the nodes' central differences (dz in
        metres).  Returns (J, grad, (forward info, adjoint info)), and with
        return_terms an (n, 4) array of each node's detector-field,
        right-hand-side, near-block and far-field shares of grad."""
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        w = np.ascontiguousarray(weights, dtype=np.float64).reshape(-1)
        if w.shape[0] != pts.shape[0]:
            raise ValueError("one weight per point")
        nd = np.ascontiguousarray(nodes, dtype=np.int32).reshape(-1)
        k = np.ascontiguousarray(k_in, dtype=np.float64)
        e = np.ascontiguousarray(e_in, dtype=np.float64)
        gp = _gmres_params(max_iterations, 0, tolerance, "ICGS", False, 0, 0.0, False, near_field,
                           inner_iterations)
        J = np.zeros(1)
        grad = np.zeros(max(nd.size, 1))
        terms = np.zeros((max(nd.size, 1), 4))
        infos = (nat.GmresInfo * 2)()
        nat.check(nat.lib().spk_adjoint_height_gradient(
            self._h, _dp(k), _dp(e), pts.shape[0], _dp(pts), _dp(w), C.byref(gp), nd.size,
            nd.ctypes.data_as(C.POINTER(C.c_int)), float(dz), _dp(J), _dp(grad), _dp(terms),
            infos))
        info = tuple(dict(info=i.info, iters=i.iterations, be=i.backward_error) for i in infos)
        if return_terms:
            return float(J[0]), grad[:nd.size], info, terms[:nd.size]
        return float(J[0]), grad[:nd.size], info

    # -- Integration_scattered_field_tri -------------------------------------
    def scattered_field(self, points, vector_x):
        """Scattered (E, H) at points (m) in the exterior medium from the
        surface currents x = [J; M]; arrays of shape (npts, 3) complex."""
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        x = np.ascontiguousarray(vector_x, dtype=np.complex128)
        if x.shape != (2 * self.N,):
            raise ValueError(f"vector_x must have length 2N = {2 * self.N}")
        E = np.zeros((pts.shape[0], 3), np.complex128)
        H = np.zeros((pts.shape[0], 3), np.complex128)
        nat.check(nat.lib().spk_scattered_field(self._h, pts.shape[0], _dp(pts),
                                                _dp(x.view(np.float64)),
                                                _dp(E.view(np.float64)), _dp(H.view(np.float64))))
        return E, H

    def scattered_field_device(self, points, d_x):
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        E = np.zeros((pts.shape[0], 3), np.complex128)
        H = np.zeros((pts.shape[0], 3), np.complex128)
        nat.check(nat.lib().spk_scattered_field_device(self._h, pts.shape[0], _dp(pts),
                                                       C.c_void_p(int(d_x)),
                                                       _dp(E.view(np.float64)),
                                                       _dp(H.view(np.float64))))
        return E, H

    # -- lib_math_solver_gmres_run ------------------------------------------
    def gmres_run(self, vector_b, x0=None, max_iterations=650, restart=0, tolerance=1e-3,
                  orthogonalization="IMGS", use_initial_guess=True, refinement_steps=0,
                  refinement_tolerance=0.0, precondition_left=False, near_field=False,
                  inner_iterations=0):
        """CERFACS GMRES on the FP64 operator (precondition_left: the
        reference's Jacobi left preconditioner); on a precision="mixed"
        handle, FP64 iterative refinement around complex64 inner GMRES
        solves (refinement_tolerance: inner relative residual of the first
        solve, default 3e-4; later solves stop at tolerance / (2 |r|/|b|)).
        near_field: flexible GMRES on the FP64 operator, right-preconditioned
        by inner_iterations (default 30) GMRES steps on the near-field
        operator (complex64 on a mixed handle); restart defaults to 100."""
        b = np.ascontiguousarray(vector_b, dtype=np.complex128)
        x = self.get_init_vector_x_tri() if x0 is None else np.array(x0, dtype=np.complex128)
        gp = _gmres_params(max_iterations, restart, tolerance, orthogonalization,
                           use_initial_guess, refinement_steps, refinement_tolerance,
                           precondition_left, near_field, inner_iterations)
        info = nat.GmresInfo()
        nat.check(nat.lib().spk_gmres(self._h, _dp(b.view(np.float64)), _dp(x.view(np.float64)),
                                      C.byref(gp), C.byref(info)))
        return x, dict(info=info.info, iters=info.iterations,
                       bea=info.arnoldi_backward_error, be=info.backward_error,
                       tbe=info.true_backward_error, matvecs=info.matvecs,
                       seconds=info.seconds, hist=self.gmres_history())

    def gmres_run_device(self, d_b, d_x, max_iterations=650, restart=0, tolerance=1e-3,
                         orthogonalization="IMGS", use_initial_guess=True, refinement_steps=0,
                         refinement_tolerance=0.0, precondition_left=False, near_field=False,
                         inner_iterations=0):
        gp = _gmres_params(max_iterations, restart, tolerance, orthogonalization,
                           use_initial_guess, refinement_steps, refinement_tolerance,
                           precondition_left, near_field, inner_iterations)
        info = nat.GmresInfo()
        nat.check(nat.lib().spk_gmres_device(self._h, C.c_void_p(int(d_b)), C.c_void_p(int(d_x)),
                                             C.byref(gp), C.byref(info)))
        return dict(info=info.info, iters=info.iterations, bea=info.arnoldi_backward_error,
                    be=info.backward_error, tbe=info.true_backward_error, matvecs=info.matvecs,
                    seconds=info.seconds)

    def gmres_run_batch_device(self, d_bs, d_xs, max_iterations=650, tolerance=1e-3,
                               use_initial_guess=True, refinement_steps=0,
                               refinement_tolerance=0.0, near_field=False, inner_iterations=0,
                               restart=0):
        """Solves of R <= 16 right-hand sides (device pointers to 2N
        complex128 vectors) sharing batched products.  Default: mixed
        precision refinement on a precision="mixed" handle, the inner
        complex64 steps sharing one batched matvec.  near_field: flexible
        GMRES on the FP64 operator preconditioned by inner_iterations GMRES
        steps on the near-field operator (any handle).  Returns one info dict
        per right-hand side."""
        R = len(d_bs)
        if len(d_xs) != R:
            raise ValueError("d_bs and d_xs must have the same length")
        gp = _gmres_params(max_iterations, restart, tolerance, "ICGS", use_initial_guess,
                           refinement_steps, refinement_tolerance, False, near_field,
                           inner_iterations)
        infos = (nat.GmresInfo * R)()
        bs = (C.c_void_p * R)(*[int(p) for p in d_bs])
        xs = (C.c_void_p * R)(*[int(p) for p in d_xs])
        nat.check(nat.lib().spk_gmres_batch_device(self._h, R, bs, xs, C.byref(gp), infos))
        return [dict(info=i.info, iters=i.iterations, bea=i.arnoldi_backward_error,
                     be=i.backward_error, tbe=i.true_backward_error, matvecs=i.matvecs,
                     seconds=i.seconds) for i in infos]

    def gmres_history(self):
        n = C.c_int()
        nat.check(nat.lib().spk_gmres_history(self._h, None, 0, C.byref(n)))
        h = np.zeros(max(n.value, 1))
        nat.check(nat.lib().spk_gmres_history(self._h, _dp(h), n.value, C.byref(n)))
        return h[:n.value]

    # -- probes ---------------------------------------------------------------
    def stats(self):
        s = nat.Stats()
        nat.check(nat.lib().spk_get_stats(self._h, C.byref(s)))
        d = {k: getattr(s, k) for k, _ in nat.Stats._fields_}
        d["boxes_per_level"] = list(s.boxes_per_level)[: s.l_max + 1]
        d["i4_pairs_per_level"] = list(s.i4_pairs_per_level)[: s.l_max + 1]
        return d

    def rwg(self):
        c = np.zeros((self.N, 2), np.int32)
        e = np.zeros((self.N, 2), np.int32)
        nat.check(nat.lib().spk_get_rwg(self._h, c.ctypes.data_as(C.POINTER(C.c_int)),
                                        e.ctypes.data_as(C.POINTER(C.c_int))))
        return c, e

    def tree(self):
        leaf = np.zeros(self.N, np.int64)
        a, b = C.c_int(), C.c_int()
        D = np.zeros(1)
        xm = np.zeros(3)
        nat.check(nat.lib().spk_get_tree(self._h, leaf.ctypes.data_as(C.POINTER(C.c_int64)),
                                         C.byref(a), C.byref(b), _dp(D), _dp(xm)))
        return dict(l_min=a.value, l_max=b.value, D=D[0], xmin=xm, leaf=leaf)

    def i4(self, level, n):
        cnt = C.c_int()
        out = np.zeros(216, np.int64)
        nat.check(nat.lib().spk_get_i4(self._h, level, n, out.ctypes.data_as(C.POINTER(C.c_int64)),
                                       216, C.byref(cnt)))
        return out[:cnt.value]

    def near_csr(self, values=True):
        st = self.stats()
        nnz = st["near_blocks"]
        ia = np.zeros(self.N + 1, np.int64)
        ja = np.zeros(nnz, np.int32)
        v = np.zeros(nnz * 8) if values else None
        nat.check(nat.lib().spk_get_near_csr(self._h, ia.ctypes.data_as(C.POINTER(C.c_int64)),
                                             ja.ctypes.data_as(C.POINTER(C.c_int32)),
                                             _dp(v) if values else None))
        vals = (v[0::2] + 1j * v[1::2]).reshape(nnz, 4) if values else None
        return ia, ja, vals

    def impedance_edges(self, m, n):
        m = np.ascontiguousarray(np.atleast_1d(m), dtype=np.int32)
        n = np.ascontiguousarray(np.atleast_1d(n), dtype=np.int32)
        out = np.zeros(8 * m.size)
        nat.check(nat.lib().spk_near_blocks(self._h, m.size, m.ctypes.data_as(C.POINTER(C.c_int)),
                                            n.ctypes.data_as(C.POINTER(C.c_int)), _dp(out)))
        return (out[0::2] + 1j * out[1::2]).reshape(-1, 2, 2)

    def profile_near_bf(self, reps=20, fp8=False):
        """Mean device time [ms] and algorithmic HBM bytes of the near-field
        preconditioner's SpMV (bf16 blocks, or the fp8 copy with fp8=True;
        precision="mixed" handle)."""
        ms, by = np.zeros(1), np.zeros(1)
        fn = nat.lib().spk_profile_near_f8 if fp8 else nat.lib().spk_profile_near_bf
        nat.check(fn(self._h, reps, _dp(ms), _dp(by)))
        return float(ms[0]), float(by[0])

    def near_product_lowp(self, X, fmt="fp8", batched=False):
        """The near-field part of a precision="mixed" handle's operator on the
        rows of X (R <= 16 complex64 vectors of length 2N), with the blocks in
        complex64 ("c64") or in the preconditioner's "bf16" / "fp8" copy;
        batched: the batched launch of the inner solves instead of one
        single-vector launch per vector."""
        from ._native import DeviceBuffer

        X = np.ascontiguousarray(np.atleast_2d(X), dtype=np.complex64)
        R = X.shape[0]
        nb = 2 * self.N * 8
        xb = [DeviceBuffer(nb).upload(X[r]) for r in range(R)]
        yb = [DeviceBuffer(nb) for _ in range(R)]
        xs = (C.c_void_p * R)(*[b.ptr for b in xb])
        ys = (C.c_void_p * R)(*[b.ptr for b in yb])
        f = {"c64": 0, "bf16": 1, "fp8": 2}[fmt]
        nat.check(nat.lib().spk_near_product_lowp(self._h, f, int(bool(batched)), R, xs, ys))
        return np.stack([yb[r].download(np.complex64) for r in range(R)])

    def profile_matvec(self, reps=20, c64=False):
        """Mean device time [ms] and algorithmic HBM bytes of each matvec
        stage, measured with HIP events on the handle's stream (c64: the
        complex64 operator of a mixed handle)."""
        ms = np.zeros(8)
        by = np.zeros(8)
        f = nat.lib().spk_profile_matvec_c64 if c64 else nat.lib().spk_profile_matvec
        nat.check(f(self._h, reps, _dp(ms), _dp(by)))
        return {k: (float(ms[i]), float(by[i])) for i, k in enumerate(nat.STAGES)}

    def patterns(self, first, count):
        S = self.stats()["samples"]
        out = np.zeros(count * 8 * S)
        nat.check(nat.lib().spk_get_patterns(self._h, first, count, _dp(out)))
        return (out[0::2] + 1j * out[1::2]).reshape(count, 4, S)


def ensemble_intensity(wavelength, eps_r2, meshes, angles, points, eps_r1=1.0 + 0.0j,
                       formulation="ICTF", normal_convention=2, tree_s_opt=250,
                       truncation_number=19, precision="mixed", device=0, tolerance=1e-6,
                       max_iterations=650, inner_iterations=0, workers=3):
    """A speckle ensemble (spk_ensemble_intensity): every mesh (pp, tt) in
    ``meshes`` is solved for every plane wave (k_in, e_in) in ``angles`` by
    the near-field preconditioned flexible GMRES, in batches of up to 16, and
    the scattered intensity sum |E_c|^2 is evaluated at ``points`` (npts, 3).
    ``workers`` host threads with their own handles take the meshes in turn.
    ``inner_iterations="tune"`` tunes on the first mesh first (spk_autotune)
    and applies the result to every worker.
    Returns (intensity (n_mesh, n_angles, npts), info dicts (n_mesh, n_angles))."""
    nat.check_single_hip_runtime()
    f = FORMULATIONS[formulation] if isinstance(formulation, str) else int(formulation)
    prec = {"fp64": 0, "mixed": 1}[precision]
    eps_r1, eps_r2 = complex(eps_r1), complex(eps_r2)
    params = nat.Params(wavelength, eps_r1.real, eps_r1.imag, eps_r2.real, eps_r2.imag, f,
                        normal_convention, 3, tree_s_opt, truncation_number, prec, device)
    keep = []
    ms = (nat.Mesh * len(meshes))()
    for i, (pp, tt) in enumerate(meshes):
        pp = np.ascontiguousarray(pp, dtype=np.float64).reshape(-1, 3)
        tt = np.ascontiguousarray(tt, dtype=np.int32).reshape(-1, 3)
        keep += [pp, tt]
        ms[i] = nat.Mesh(pp.shape[0], _dp(pp), tt.shape[0], tt.ctypes.data_as(C.POINTER(C.c_int)))
    kin = np.ascontiguousarray([a[0] for a in angles], dtype=np.float64).reshape(-1, 3)
    ein = np.ascontiguousarray([a[1] for a in angles], dtype=np.float64).reshape(-1, 3)
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    nm, na, npts = len(meshes), kin.shape[0], pts.shape[0]
    out = np.zeros((nm, na, npts))
    infos = (nat.GmresInfo * max(1, nm * na))()
    gp = _gmres_params(max_iterations, 0, tolerance, "ICGS", False, 0, 0.0, False, True,
                       -1 if inner_iterations == "tune" else inner_iterations)
    nat.check(nat.lib().spk_ensemble_intensity(C.byref(params), nm, ms, na, _dp(kin), _dp(ein),
                                               npts, _dp(pts), C.byref(gp), int(workers),
                                               _dp(out), infos))
    info = [[dict(info=infos[m * na + a].info, iters=infos[m * na + a].iterations,
                  be=infos[m * na + a].backward_error) for a in range(na)] for m in range(nm)]
    return out, info
