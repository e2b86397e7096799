"""Synthetic meshes and the reference's mesh file formats.

* ``icosphere`` -- flat-triangle sphere (icosahedron with midpoint
  subdivision projected to the sphere): 20*4**level triangles.
* ``octasphere`` -- the reference's own recursive octahedral sphere
  (``recursive_triangulation``, src/lib_sie_tri_calculation_mod.f90:646-760),
  node order and corner order kept.
* ``rough_surface`` -- Gaussian random height field on a regular grid
  (counter-based SplitMix64 + Box-Muller, Gaussian spectral filter), each cell
  split into two triangles along one diagonal.
* ``read_pp_tt`` / ``write_pp_tt`` -- the ``pp_*``/``tt_*`` text files of
  ``read_pp_tt_file_tri`` (src/lib_sie_tri_data_container.f90:226-261): one
  node (x y z) or one triangle (three 1-based node indices) per line.

All triangle index arrays returned here are 1-based, as the reference reads
them.
"""
from __future__ import annotations

import numpy as np

_MASK64 = (1 << 64) - 1


def _splitmix64(seed: int, n: int) -> np.ndarray:
    """n outputs of SplitMix64 (Steele, Lea, Flood 2014) as uint64."""
    out = np.empty(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = (np.uint64(seed & _MASK64) + idx * np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    out[:] = z
    return out


def gaussian_normals(seed: int, n: int) -> np.ndarray:
    """n standard normal deviates by Box-Muller on SplitMix64 uniforms."""
    m = (n + 1) // 2
    u = _splitmix64(seed, 2 * m)
    # 53-bit uniforms in (0, 1]
    f = ((u >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)
    u1, u2 = f[0::2], f[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
    return z[:n]


def rough_surface(nx: int = 100, ny: int = 100, lx: float = 10.0, ly: float = 10.0,
                  rms: float = 0.1, corr: float = 0.5, seed: int = 1,
                  wavelength: float = 633e-9):
    """Gaussian rough surface patch of nx*ny cells (lengths in wavelengths).

    Heights have a Gaussian correlation function exp(-r^2/corr^2) and RMS
    ``rms``; the filter is applied in the spectral domain on the periodic
    (nx+1) x (ny+1) node grid.  Returns (pp [m], tt 1-based).
    """
    gx, gy = nx + 1, ny + 1
    w = gaussian_normals(seed, gx * gy).reshape(gy, gx)
    kx = 2 * np.pi * np.fft.fftfreq(gx, d=lx / nx)
    ky = 2 * np.pi * np.fft.fftfreq(gy, d=ly / ny)
    kk = kx[None, :] ** 2 + ky[:, None] ** 2
    filt = np.exp(-kk * corr * corr / 8.0)  # sqrt of the Gaussian spectrum
    h = np.real(np.fft.ifft2(np.fft.fft2(w) * filt))
    h = h - h.mean()
    h = h * (rms / np.sqrt(np.mean(h * h)))
    xs = np.linspace(-lx / 2, lx / 2, gx)
    ys = np.linspace(-ly / 2, ly / 2, gy)
    X, Y = np.meshgrid(xs, ys)
    pp = np.stack([X.ravel(), Y.ravel(), h.ravel()], axis=1) * wavelength
    tt = []
    for j in range(ny):
        for i in range(nx):
            a = j * gx + i + 1
            b = a + 1
            c = a + gx
            d = c + 1
            tt.append((a, b, d))  # counter-clockwise seen from +z
            tt.append((a, d, c))
    return pp, np.asarray(tt, dtype=np.int32)


def icosphere(level: int = 3, radius: float = 1.0):
    """Icosahedron subdivided ``level`` times, nodes projected to the sphere.

    Triangles are oriented counter-clockwise seen from outside.  Returns
    (pp, tt 1-based): 20*4**level triangles, 10*4**level+2 nodes.
    """
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
         (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.asarray(p, float) / np.linalg.norm(p) for p in v]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                p = verts[a] + verts[b]
                verts.append(p / np.linalg.norm(p))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    pp = np.asarray(verts) * radius
    tt = np.asarray(faces, dtype=np.int32) + 1
    return pp, tt


def octasphere(n_disc: int, radius: float):
    """The reference's recursive octahedral sphere with n_disc nodes
    (n_disc = 6, 18, 66, 258, 1026, ...): recursive_triangulation and
    recursive_split (src/lib_sie_tri_calculation_mod.f90:646-760)."""
    pts = np.zeros((n_disc, 3))
    base = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    for i, p in enumerate(base):
        pts[i] = np.asarray(p, float) * radius
    ne = 2 * (n_disc - 2)
    el = [[1, 5, 3], [3, 5, 2], [2, 5, 4], [4, 5, 1], [1, 3, 6], [3, 2, 6], [2, 4, 6], [4, 1, 6]]
    el += [[0, 0, 0]] * (ne - 8)
    np_curr, ne_curr = 6, 8
    lim = int(np.log10(ne / 6.0) / np.log10(4.0))
    counter = [np_curr]  # next free slot, 0-based

    def normalize(v):
        return v / np.sqrt(v[0] ** 2 + v[1] ** 2 + v[2] ** 2)

    def point_in_list(p):
        d = np.sqrt(((pts - p) ** 2).sum(axis=1))
        hit = np.nonzero(d < 1e-30)[0]
        return int(hit[0]) + 1 if hit.size else -1

    for _ in range(lim):
        counter[0] = np_curr
        for i in range(ne_curr):
            c1, c2, c3 = el[i]
            a = normalize(0.5 * (pts[c1 - 1] + pts[c3 - 1])) * radius
            b = normalize(0.5 * (pts[c1 - 1] + pts[c2 - 1])) * radius
            c = normalize(0.5 * (pts[c2 - 1] + pts[c3 - 1])) * radius
            ids = []
            for q in (a, b, c):
                k = point_in_list(q)
                if k == -1:
                    pts[counter[0]] = q
                    counter[0] += 1
                    k = counter[0]
                ids.append(k)
            ia, ib, ic = ids
            el[ne_curr + i * 3 + 0] = [c1, ib, ia]
            el[ne_curr + i * 3 + 1] = [ib, c2, ic]
            el[ne_curr + i * 3 + 2] = [ia, ic, c3]
            el[i] = [ia, ib, ic]
        ne_curr *= 4
        np_curr = ne_curr // 2 + 2
    return pts, np.asarray(el, dtype=np.int32)


def write_pp_tt(pp, tt, pp_path, tt_path):
    np.savetxt(pp_path, np.asarray(pp), fmt="%.17e")
    np.savetxt(tt_path, np.asarray(tt, dtype=np.int64), fmt="%d")


def read_pp_tt(pp_path, tt_path):
    pp = np.loadtxt(pp_path, dtype=np.float64).reshape(-1, 3)
    tt = np.loadtxt(tt_path, dtype=np.int64).reshape(-1, 3).astype(np.int32)
    return pp, tt


# ---------------------------------------------------------------------------
# 8-node quadrilateral height fields (the reference's quad surface input)
# ---------------------------------------------------------------------------
def rough_height_field(nx: int = 10, ny: int = 10, lx: float = 2.0, ly: float = 2.0,
                       rms: float = 0.1, corr: float = 0.5, seed: int = 1,
                       wavelength: float = 633e-9):
    """Gaussian rough height field for nx x ny 8-node quadrilaterals: node
    heights on the (2nx+1) x (2ny+1) grid, same statistics and generator as
    rough_surface.  Returns (Lx, Ly, nx, ny, z) with lengths in metres and z
    in the height file's order (x index fastest)."""
    gx, gy = 2 * nx + 1, 2 * ny + 1
    w = gaussian_normals(seed, gx * gy).reshape(gy, gx)
    kx = 2 * np.pi * np.fft.fftfreq(gx, d=lx / (gx - 1))
    ky = 2 * np.pi * np.fft.fftfreq(gy, d=ly / (gy - 1))
    kk = kx[None, :] ** 2 + ky[:, None] ** 2
    h = np.real(np.fft.ifft2(np.fft.fft2(w) * np.exp(-kk * corr * corr / 8.0)))
    h = h - h.mean()
    h = h * (rms / np.sqrt(np.mean(h * h)))
    return lx * wavelength, ly * wavelength, nx, ny, (h * wavelength).ravel()


def write_quad_height_file(path, Lx, Ly, nx, ny, z):
    """The reference's quad surface file (src/lib_sie_quad_data_container.f90:330-345):
    line 1 `Lx Ly`, line 2 `Nx Ny`, then (2Nx+1)(2Ny+1) heights, x index fastest."""
    with open(path, "w") as f:
        f.write(f"{Lx:.17e} {Ly:.17e}\n{nx} {ny}\n")
        for v in np.asarray(z, float).ravel():
            f.write(f"{v:.17e}\n")


def read_quad_height_file(path):
    with open(path) as f:
        toks = f.read().split()
    Lx, Ly = float(toks[0]), float(toks[1])
    nx, ny = int(float(toks[2])), int(float(toks[3]))
    z = np.array([float(t) for t in toks[4:4 + (2 * nx + 1) * (2 * ny + 1)]])
    return Lx, Ly, nx, ny, z
