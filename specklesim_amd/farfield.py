"""Far-zone evaluations built on the scattered field: BRDF and RCS.

Mirrors the reference's evaluation modes ``calc_p(4)`` = 4..7
(src/lib_sie_tri_solver_normal.f90:325-358): the scattered field is
evaluated with the full Green's function at points on a sphere of radius
``dim_c`` in the plane phi = ``dim_b(1)``
(get_evaluation_points_domain_new, src/lib_sie_math.f90:528-535), and

* RCS:  |E|^2 * 4 pi r^2
* BRDF: |E|^2 * r^2 / (cos(theta_in) * cos(theta_s))
"""
from __future__ import annotations

import numpy as np


def sphere_points(theta_range, n_theta, phi, radius):
    """Points r = radius (sin t cos p, sin t sin p, cos t) for n_theta polar
    angles t from theta_range[0] to theta_range[1] inclusive."""
    t0, t1 = theta_range
    dx = (t1 - t0) / (n_theta - 1) if n_theta > 1 else 0.0
    theta = t0 + dx * np.arange(n_theta)
    pts = np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)],
                   axis=1) * radius
    return theta, pts


def intensity(E, points):
    """|E|^2 r^2 per point (dot_product(E, E) * vec_len(r)**2)."""
    r2 = np.sum(np.asarray(points) ** 2, axis=1)
    return np.sum(np.abs(E) ** 2, axis=1).real * r2


def brdf(solver, vector_x, theta_in, theta_range, n_theta, phi, radius, device=False):
    """BRDF curve of a solution: (theta_s, values)."""
    theta, pts = sphere_points(theta_range, n_theta, phi, radius)
    if device:
        E, _ = solver.scattered_field_device(pts, vector_x)
    else:
        E, _ = solver.scattered_field(pts, vector_x)
    return theta, intensity(E, pts) / (np.cos(theta_in) * np.cos(theta))


def rcs(solver, vector_x, theta_range, n_theta, phi, radius, device=False):
    theta, pts = sphere_points(theta_range, n_theta, phi, radius)
    if device:
        E, _ = solver.scattered_field_device(pts, vector_x)
    else:
        E, _ = solver.scattered_field(pts, vector_x)
    return theta, intensity(E, pts) * 4 * np.pi
