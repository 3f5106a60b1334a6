"""MI355X-native sparse-block-grid smoother and particle-transfer path.

Host-side mirror of the msbg-rs API for this path (SparseGrid, Sweeper,
PdeParams, Stencil, MacGrid, particles_to_grid, MacSampler, ...) over the
C-ABI library `libmsbg_hip.so` (HIP kernels for gfx950).
"""
from .grid import SparseGrid, morton3, EMPTY_SLOT, FULL_SLOT  # noqa: F401
from . import lib  # noqa: F401
