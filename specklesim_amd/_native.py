"""ctypes binding of libspeckle_gpu.so (the C-ABI in include/speckle_gpu.h).

The library is the product: there is no fallback.  If it is missing or the
GPU is unavailable, calls raise ``SpeckleError``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspeckle_gpu.so")

# Exported entry points, in include/speckle_gpu.h order.
EXPORTS = [
    "spk_create", "spk_destroy", "spk_last_error", "spk_setup_tri", "spk_setup_quad", "spk_get_stats",
    "spk_num_edges", "spk_matvec", "spk_matvec_device", "spk_set_transposed",
    "spk_get_transposed", "spk_set_inner_graph", "spk_adjoint_height_gradient",
    "spk_autotune", "spk_set_tuning", "spk_set_level_truncation", "spk_interp_probe",
    "spk_matvec_c64",
    "spk_matvec_c64_device", "spk_matvec_parts", "spk_rhs_plane_wave",
    "spk_rhs_plane_wave_device", "spk_scattered_field", "spk_scattered_field_device",
    "spk_scattered_field_device_noalloc",
    "spk_quad_create", "spk_quad_destroy", "spk_quad_setup", "spk_quad_sizes",
    "spk_quad_get_mesh", "spk_quad_get_edges", "spk_quad_edge_element", "spk_quad_fill",
    "spk_quad_fill_device", "spk_quad_topology_probe", "spk_quad_rhs_plane_wave",
    "spk_quad_rhs_plane_wave_device", "spk_quad_solve_lu_device", "spk_quad_solve_plane_wave",
    "spk_quad_scattered_field", "spk_probe_read_bandwidth", "spk_matvec_batch_device",
    "spk_matvec_batch_c64_device", "spk_ensemble_intensity", "spk_gmres_batch_device",
    "spk_device_alloc", "spk_device_free", "spk_device_copy", "spk_device_synchronize",
    "spk_gmres", "spk_gmres_device", "spk_gmres_history", "spk_get_rwg", "spk_get_tree",
    "spk_get_i4", "spk_get_near_csr", "spk_near_blocks", "spk_get_patterns",
    "spk_profile_matvec", "spk_profile_matvec_c64", "spk_profile_near_bf", "spk_profile_near_f8",
    "spk_topology_probe", "spk_near_product_lowp",
]
STAGES = ["aggregation", "ss", "m2l", "rr", "fold", "disaggregation", "near_spmv", "matvec"]


class SpeckleError(RuntimeError):
    pass


class Params(C.Structure):
    _fields_ = [("lambda_", C.c_double), ("eps_r1_re", C.c_double), ("eps_r1_im", C.c_double),
                ("eps_r2_re", C.c_double), ("eps_r2_im", C.c_double), ("formulation", C.c_int),
                ("normal_convention", C.c_int), ("quad_points", C.c_int),
                ("tree_s_opt", C.c_int), ("truncation_number", C.c_int),
                ("precision", C.c_int), ("device", C.c_int)]


class GmresParams(C.Structure):
    _fields_ = [("max_iterations", C.c_int), ("restart", C.c_int), ("tolerance", C.c_double),
                ("orthogonalization", C.c_int), ("use_initial_guess", C.c_int),
                ("refinement_steps", C.c_int), ("refinement_tolerance", C.c_double),
                ("preconditioning", C.c_int), ("inner_iterations", C.c_int)]


class GmresInfo(C.Structure):
    _fields_ = [("info", C.c_int), ("iterations", C.c_int),
                ("arnoldi_backward_error", C.c_double), ("backward_error", C.c_double),
                ("matvecs", C.c_int), ("seconds", C.c_double),
                ("true_backward_error", C.c_double)]


class Mesh(C.Structure):
    _fields_ = [("np", C.c_int), ("pp", C.POINTER(C.c_double)), ("nt", C.c_int),
                ("tt", C.POINTER(C.c_int))]


class Tuning(C.Structure):
    _fields_ = [("tiles_per_group", C.c_int), ("inner_iterations", C.c_int),
                ("tile_us", C.c_double * 3), ("n_inner", C.c_int), ("inner_tried", C.c_int * 8),
                ("inner_ms", C.c_double * 8)]


class Stats(C.Structure):
    _fields_ = [("n_edges", C.c_int64), ("near_blocks", C.c_int64), ("l_min", C.c_int),
                ("l_max", C.c_int), ("samples", C.c_int),
                ("boxes_per_level", C.c_int64 * 32), ("i4_pairs_per_level", C.c_int64 * 32),
                ("singular_pairs", C.c_int64), ("setup_ms", C.c_double),
                ("near_fill_ms", C.c_double), ("patterns_ms", C.c_double),
                ("tables_ms", C.c_double), ("tree_ms", C.c_double), ("bytes_near", C.c_double),
                ("bytes_patterns", C.c_double)]


_lib = None


def hip_runtimes(maps_text: str) -> list:
    """The distinct HIP runtime files (libamdhip64.so*) mapped into a process,
    from the text of its /proc/<pid>/maps: one entry per (device, inode),
    named by the first path seen for it."""
    seen = {}
    for line in maps_text.splitlines():
        parts = line.split(None, 5)
        if len(parts) < 6:
            continue
        path = parts[5].strip()
        if not os.path.basename(path).startswith("libamdhip64.so"):
            continue
        seen.setdefault((parts[3], parts[4]), path)
    return sorted(seen.values())


def check_single_hip_runtime(maps_text: str | None = None) -> None:
    """Raise SpeckleError when two HIP runtimes are mapped into this process.

    PyTorch wheels carry their own libamdhip64.so, whose soname equals the
    ROCm one but whose file name does not: when this library is loaded
    first, a later ``import torch`` maps a second runtime.  Two runtimes own
    two device contexts; pointers from one are invalid in the other, and the
    pair has been seen to crash at process exit.  Importing torch before
    this library keeps one runtime (the loader then matches the soname)."""
    if maps_text is None:
        try:
            with open("/proc/self/maps") as f:
                maps_text = f.read()
        except OSError:
            return
    rts = hip_runtimes(maps_text)
    if len(rts) > 1:
        raise SpeckleError("two HIP runtimes are loaded in this process (" + ", ".join(rts) +
                           "): import torch before specklesim_amd, or do not import torch")


def lib():
    """Load libspeckle_gpu.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SpeckleError(f"{LIB_PATH} is missing: build it with "
                           "`python -m specklesim_amd.build` (no CPU fallback exists)")
    L = C.CDLL(LIB_PATH)
    vp, dp, ip, i64p = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int64)
    L.spk_last_error.restype = C.c_char_p
    L.spk_create.argtypes = [C.POINTER(Params), C.POINTER(vp)]
    L.spk_destroy.argtypes = [vp]
    L.spk_destroy.restype = None
    L.spk_setup_tri.argtypes = [vp, C.c_int, dp, C.c_int, ip]
    L.spk_setup_quad.argtypes = [vp, C.c_double, C.c_double, C.c_int, C.c_int, dp]
    L.spk_get_stats.argtypes = [vp, C.POINTER(Stats)]
    L.spk_num_edges.argtypes = [vp, i64p]
    L.spk_matvec.argtypes = [vp, dp, dp]
    L.spk_matvec_device.argtypes = [vp, vp, vp]
    L.spk_matvec_c64.argtypes = [vp, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.spk_matvec_c64_device.argtypes = [vp, vp, vp]
    L.spk_matvec_parts.argtypes = [vp, dp, dp, dp]
    L.spk_rhs_plane_wave.argtypes = [vp, dp, dp, dp]
    L.spk_rhs_plane_wave_device.argtypes = [vp, dp, dp, vp]
    L.spk_scattered_field.argtypes = [vp, C.c_int, dp, dp, dp, dp]
    L.spk_quad_create.argtypes = [C.POINTER(Params), C.POINTER(vp)]
    L.spk_quad_destroy.argtypes = [vp]
    L.spk_quad_destroy.restype = None
    L.spk_quad_setup.argtypes = [vp, C.c_double, C.c_double, C.c_int, C.c_int, dp]
    L.spk_quad_sizes.argtypes = [vp, ip, ip, ip]
    L.spk_quad_get_mesh.argtypes = [vp, dp, ip]
    L.spk_quad_get_edges.argtypes = [vp, ip]
    L.spk_quad_edge_element.argtypes = [vp, C.c_int, ip, ip, dp]
    L.spk_quad_fill.argtypes = [vp, dp]
    L.spk_quad_fill_device.argtypes = [vp, vp]
    L.spk_quad_rhs_plane_wave.argtypes = [vp, dp, dp, dp]
    L.spk_quad_rhs_plane_wave_device.argtypes = [vp, dp, dp, vp]
    L.spk_quad_solve_lu_device.argtypes = [vp, vp, C.c_int, vp]
    L.spk_quad_solve_plane_wave.argtypes = [vp, dp, dp, dp]
    L.spk_quad_scattered_field.argtypes = [vp, C.c_int, dp, dp, dp, dp]
    L.spk_probe_read_bandwidth.argtypes = [C.c_int, C.c_size_t, C.c_int, dp]
    vpp = C.POINTER(C.c_void_p)
    L.spk_matvec_batch_device.argtypes = [vp, C.c_int, vpp, vpp]
    L.spk_matvec_batch_c64_device.argtypes = [vp, C.c_int, vpp, vpp]
    L.spk_near_product_lowp.argtypes = [vp, C.c_int, C.c_int, C.c_int, vpp, vpp]
    L.spk_set_transposed.argtypes = [vp, C.c_int]
    L.spk_set_inner_graph.argtypes = [vp, C.c_int]
    L.spk_autotune.argtypes = [vp, C.c_int, dp, dp, C.POINTER(GmresParams), C.POINTER(Tuning)]
    L.spk_set_tuning.argtypes = [vp, C.c_int, C.c_int]
    L.spk_set_level_truncation.argtypes = [vp, C.c_int, ip, C.c_int]
    L.spk_interp_probe.argtypes = [C.c_int, C.c_int, dp, dp]
    L.spk_adjoint_height_gradient.argtypes = [vp, dp, dp, C.c_int, dp, dp, C.POINTER(GmresParams),
                                              C.c_int, ip, C.c_double, dp, dp, dp,
                                              C.POINTER(GmresInfo)]
    L.spk_get_transposed.argtypes = [vp, C.POINTER(C.c_int)]
    L.spk_gmres_batch_device.argtypes = [vp, C.c_int, vpp, vpp, C.POINTER(GmresParams),
                                         C.POINTER(GmresInfo)]
    L.spk_ensemble_intensity.argtypes = [C.POINTER(Params), C.c_int, C.POINTER(Mesh), C.c_int,
                                         dp, dp, C.c_int, dp, C.POINTER(GmresParams), C.c_int,
                                         dp, C.POINTER(GmresInfo)]
    L.spk_quad_topology_probe.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, dp, ip, ip, ip,
                                          dp, ip, ip]
    L.spk_device_alloc.argtypes = [C.c_int, C.c_size_t, C.POINTER(vp)]
    L.spk_device_free.argtypes = [vp]
    L.spk_device_copy.argtypes = [vp, vp, C.c_size_t, C.c_int]
    L.spk_device_synchronize.argtypes = [C.c_int]
    L.spk_scattered_field_device.argtypes = [vp, C.c_int, dp, vp, dp, dp]
    L.spk_scattered_field_device_noalloc.argtypes = [vp, C.c_int, dp, vp, dp, dp]
    L.spk_gmres.argtypes = [vp, dp, dp, C.POINTER(GmresParams), C.POINTER(GmresInfo)]
    L.spk_gmres_device.argtypes = [vp, vp, vp, C.POINTER(GmresParams), C.POINTER(GmresInfo)]
    L.spk_gmres_history.argtypes = [vp, dp, C.c_int, ip]
    L.spk_get_rwg.argtypes = [vp, ip, ip]
    L.spk_get_tree.argtypes = [vp, i64p, ip, ip, dp, dp]
    L.spk_get_i4.argtypes = [vp, C.c_int, C.c_int64, i64p, C.c_int, ip]
    L.spk_get_near_csr.argtypes = [vp, i64p, C.POINTER(C.c_int32), dp]
    L.spk_near_blocks.argtypes = [vp, C.c_int, ip, ip, dp]
    L.spk_get_patterns.argtypes = [vp, C.c_int, C.c_int, dp]
    L.spk_profile_matvec.argtypes = [vp, C.c_int, dp, dp]
    L.spk_profile_matvec_c64.argtypes = [vp, C.c_int, dp, dp]
    L.spk_profile_near_f8.argtypes = [vp, C.c_int, dp, dp]
    L.spk_profile_near_bf.argtypes = [vp, C.c_int, dp, dp]
    L.spk_topology_probe.argtypes = [C.c_int, dp, C.c_int, ip, C.c_int, C.c_int, i64p, ip, ip,
                                     i64p, ip, ip, dp, dp]
    check_single_hip_runtime()
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().spk_last_error()
        raise SpeckleError(f"libspeckle_gpu error {rc}: {msg.decode() if msg else ''}")


class DeviceBuffer:
    """HBM buffer owned by libspeckle_gpu's HIP runtime (for the *_device
    entry points); ``ptr`` is the device address."""

    def __init__(self, nbytes: int, device: int = 0):
        p = C.c_void_p()
        check(lib().spk_device_alloc(device, nbytes, C.byref(p)))
        self.ptr = p.value or 0
        self.nbytes = nbytes

    def upload(self, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        check(lib().spk_device_copy(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p),
                                    self.nbytes, 1))
        return self

    def download(self, dtype=np.complex128) -> np.ndarray:
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype)
        check(lib().spk_device_copy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr),
                                    self.nbytes, 2))
        return out

    def free(self) -> None:
        if self.ptr:
            lib().spk_device_free(C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize(device: int = 0) -> None:
    check(lib().spk_device_synchronize(device))


def read_bandwidth(device=0, mib=2048, reps=5):
    """Streaming-read bandwidth of the device in GB/s (spk_probe_read_bandwidth)."""
    import numpy as np

    out = np.zeros(1)
    check(lib().spk_probe_read_bandwidth(device, mib << 20, reps,
                                         out.ctypes.data_as(C.POINTER(C.c_double))))
    return float(out[0])
