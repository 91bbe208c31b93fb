"""mpifield_amd: the particle-grid hot path of mpifield (B-spline gather and scatter, the
charge-conserving path deposit, halo fill and fold) on one AMD Instinct MI355X (gfx950).

The names below are the reference's (src/mpifield/__init__.py:18-59) for the parts on this
path; they run hand-written HIP kernels (libmpfhip.so, include/mpf_hip.h) through ctypes,
with torch tensors as the device buffers.
"""

from mpifield_amd._types import BlockGrid, Bounds
from mpifield_amd.edges import CUBIC_EDGES, QUADRATIC_EDGES, WHITNEY, EdgeWeighting
from mpifield_amd.field import Field
from mpifield_amd.kernels import deposit_mode, set_deposit_mode
from mpifield_amd.splines import CLAMPED_SPLINES, SPLINES, Splines, splines
from mpifield_amd.weighting import CUBIC, LINEAR, QUADRATIC, Weighting

__version__ = "0.1.0"

__all__ = [
    "CLAMPED_SPLINES",
    "CUBIC",
    "CUBIC_EDGES",
    "LINEAR",
    "QUADRATIC",
    "QUADRATIC_EDGES",
    "SPLINES",
    "WHITNEY",
    "BlockGrid",
    "Bounds",
    "EdgeWeighting",
    "Field",
    "Splines",
    "Weighting",
    "__version__",
    "deposit_mode",
    "set_deposit_mode",
    "splines",
]
