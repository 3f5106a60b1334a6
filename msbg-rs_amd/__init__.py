"""MI355X-native sparse-block-grid smoother and particle-transfer path.

Host-side mirror of the msbg-rs API for this path (SparseGrid, Sweeper,
PdeParams, Stencil, halo fills, multires transfers, particle transfer) over
the C-ABI library `libmsbg_hip.so` (HIP kernels for gfx950).  Nothing here
computes on the CPU: every device entry point goes through the library and
raises if it is not built.
"""
from .grid import SparseGrid, morton3, EMPTY_SLOT, FULL_SLOT  # noqa: F401
from . import lib  # noqa: F401
from .lib import MsbgError, DeviceArray  # noqa: F401
from .device import (  # noqa: F401
    BoundaryCondition, DeviceGrid, PdeParams, Stencil, SweepPlan, Sweeper, apply_pde, color_buckets,
    downsample_channel_avg, halo_fill, halo_fill_multires, halo_topology, regularize_refinement_map, sweep_multires,
    sweep_levels,
    compute_block_topology, Topology,
)
from .solve import PoissonBand, Storage  # noqa: F401
from .learned import CorrectorNet  # noqa: F401
from .surface import SurfaceConfig, reconstruct_surface, reconstruct_and_smooth  # noqa: F401
