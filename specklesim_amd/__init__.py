"""MI355X-native SIE solve loop (near-field fill + MLFMM matvec + GMRES)
for SpeckleSim's triangle discretisation.  The compute path is
libspeckle_gpu.so (HIP, gfx950); see include/speckle_gpu.h."""
from .sie import SieTriMlfmm, ensemble_intensity, plane_wave, FORMULATIONS  # noqa: F401
from . import meshes, results  # noqa: F401
from .quad import QuadFill  # noqa: F401
from ._native import DeviceBuffer, synchronize  # noqa: F401
