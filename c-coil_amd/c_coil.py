"""c_coil: the reference's Python module surface (tensor, coil, coil_group submodules), backed by the gfx950
library. The native module is _c_coil; this wrapper first loads the HIP runtime shared with torch."""
import hip_runtime as _hip_runtime

_hip_runtime.preload()

from _c_coil import __version__, coil, coil_group, tensor  # noqa: E402,F401
