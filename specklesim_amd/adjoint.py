"""Far-field intensity of a rough surface as a function of its node heights,
with the adjoint gradient (spk_adjoint_height_gradient), and a
torch.autograd.Function around it for checks against finite differences.

The detector window: directions within +-half_width of the specular
direction in two orthogonal planes, on a sphere of radius ``radius``; J is
the sum of |E|^2 over the window's points (unit weights).
"""
from __future__ import annotations

import numpy as np

from .sie import SieTriMlfmm, plane_wave


def detector_window(theta_in, half_width_deg=5.0, n=5, radius=1.0):
    """n x n points around the specular direction of a wave incident at
    polar angle theta_in (phi = 0, travelling towards -z): angles
    -half_width..half_width in the plane of incidence and across it."""
    a = np.radians(np.linspace(-half_width_deg, half_width_deg, n))
    pts = []
    for ax in a:
        for ay in a:
            th = theta_in + ax
            d = np.array([np.sin(th) * np.cos(ay), np.sin(ay), np.cos(th) * np.cos(ay)])
            pts.append(radius * d / np.linalg.norm(d))
    return np.array(pts)


class HeightIntensity:
    """J(z) for a height field: the mesh (pp, tt) with pp[:, 2] replaced by
    z, solved for the plane wave (k_in, e_in); ``value`` runs a forward
    solve, ``gradient`` the adjoint."""

    def __init__(self, solver: SieTriMlfmm, pp, tt, k_in, e_in, points, weights=None,
                 tolerance=1e-10, dz=1e-6):
        self.s, self.pp, self.tt = solver, np.array(pp, float), np.asarray(tt)
        self.k_in, self.e_in = np.asarray(k_in, float), np.asarray(e_in, float)
        self.points = np.asarray(points, float)
        self.weights = np.ones(len(self.points)) if weights is None else np.asarray(weights, float)
        self.tolerance, self.dz = tolerance, dz

    def _mesh(self, z):
        pp = self.pp.copy()
        pp[:, 2] = z
        return pp

    def value(self, z) -> float:
        """J(z) by a forward solve only (the adjoint's forward solve: the
        near-field preconditioned flexible GMRES to the same tolerance)."""
        from ._native import DeviceBuffer

        s = self.s.precalculation_tri(self._mesh(z), self.tt)
        nb = 2 * s.N * 16
        db, dx = DeviceBuffer(nb), DeviceBuffer(nb)
        s.get_vector_b_tri_device(self.k_in, self.e_in, db.ptr)
        info = s.gmres_run_device(db.ptr, dx.ptr, tolerance=self.tolerance, max_iterations=650,
                                  orthogonalization="ICGS", use_initial_guess=False,
                                  near_field=True)
        if info["info"] != 0:
            raise RuntimeError(f"forward solve did not converge: {info}")
        E, _ = s.scattered_field_device(self.points, dx.ptr)
        return float(np.sum(self.weights[:, None] * np.abs(E) ** 2))

    def gradient(self, z, nodes=None):
        s = self.s.precalculation_tri(self._mesh(z), self.tt)
        nodes = np.arange(len(z)) if nodes is None else np.asarray(nodes)
        return s.adjoint_height_gradient(self.k_in, self.e_in, self.points, self.weights, nodes,
                                         dz=self.dz, tolerance=self.tolerance)


def torch_function(problem: HeightIntensity):
    """A torch.autograd.Function applying J to a float64 tensor of node
    heights, whose backward is the adjoint gradient (for checks only)."""
    import torch

    class _J(torch.autograd.Function):
        @staticmethod
        def forward(ctx, z):
            zn = z.detach().cpu().numpy().astype(np.float64)
            J, g, _ = problem.gradient(zn)
            ctx.save_for_backward(torch.as_tensor(g, dtype=z.dtype))
            return z.new_tensor(J)

        @staticmethod
        def backward(ctx, gout):
            (g,) = ctx.saved_tensors
            return gout * g.to(gout.device)

    return _J.apply


__all__ = ["detector_window", "HeightIntensity", "torch_function", "plane_wave"]
