"""Surface reconstruction from a point cloud on the GPU: the reference's
`reconstruct_surface` / `reconstruct_and_smooth` pipeline
(particles/mod.rs:214-302) without the PLY reader (which is file IO): the
caller passes the cloud's vertices.

    place (sort.rs:31-143)            msbg_surface_place + msbg_active_blocks
    fill active full + min-splat      msbg_density_splat (splat.rs:92-319)
    finalize (finalize.rs:35-71)      msbg_density_finalize
    mean-curvature smoothing          msbg_sweep_density (solver.rs, Density)

Every step runs in the HIP library; nothing computes on the host except the
vertex bounding box, as the reference's PLY loader computes it.
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import lib as _lib
from .lib import DeviceArray
from .particles import DensityGrid, active_blocks

DEMO_R_PARTICLE = 2.0  # particles/mod.rs:51
DEMO_NB_DIST = 2.0     # particles/mod.rs:52
MEAN_CURVATURE = 1     # Stencil::MeanCurvature


@dataclass
class SurfaceConfig:
    """particles/mod.rs:115-163."""
    sx: int
    sy: int
    sz: int
    n_instances: int = 0          # 0: one instance per base point (the demo's nBasePoints)
    instance_scale_factor: float = 0.01
    r_particle: float = DEMO_R_PARTICLE
    nb_dist: float = DEMO_NB_DIST
    n_smooth_iters: int = 5
    smooth_dt: float = 0.05

    @classmethod
    def demo_case1(cls, res):
        return cls(res, res, res)

    @classmethod
    def demo_case2(cls, res):
        return cls(res, res, res, instance_scale_factor=0.005, smooth_dt=0.1)

    def r_scan(self):
        return float(np.float32(self.r_particle) + np.float32(self.nb_dist))


@dataclass
class Placed:
    """particles/mod.rs:166-170: device positions (n x 3) and centre block ids,
    and the sorted footprint block set (host int64)."""
    positions: DeviceArray
    bids: DeviceArray
    active: np.ndarray


def bounding_box(vertices):
    """(bbox_min, span_max) of a vertex array, as the PLY loader reports them."""
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    lo, hi = v.min(axis=0), v.max(axis=0)
    return lo, float(np.max(hi - lo))


def place(base, bbox_min, span_max, cfg: SurfaceConfig, lib=None) -> Placed:
    """sort::place (particles/sort.rs:31-143) for a base cloud (host or device)."""
    lib = lib or _lib.load()
    b = base if isinstance(base, DeviceArray) else lib.to_device(np.ascontiguousarray(base, np.float32).reshape(-1, 3))
    n_base = b.shape[0]
    bmin = np.ascontiguousarray(bbox_min, np.float32)
    n = C.c_int64()
    args = (b.ptr, n_base, bmin.ctypes.data, span_max, cfg.sx, cfg.sy, cfg.sz, cfg.n_instances,
            cfg.instance_scale_factor)
    lib.call("msbg_surface_place", *args, None, None, 0, C.byref(n))  # count
    total = n.value
    pos = lib.empty((total, 3), np.float32)
    bids = lib.empty(total, np.int32)
    if total:
        lib.call("msbg_surface_place", *args, pos.ptr, bids.ptr, total, C.byref(n))
    act = active_blocks(cfg.sx, cfg.sy, cfg.sz, pos, cfg.r_scan(), lib=lib) if total else np.zeros(0, np.int64)
    return Placed(pos, bids, act)


def finalize(grid: DensityGrid, active, cfg: SurfaceConfig):
    """finalize::finalize (particles/finalize.rs:35-71), in place."""
    a = np.ascontiguousarray(active, np.int64)
    grid.lib.call("msbg_density_finalize", grid.layout.h, grid.slab.ptr, grid.slots, a.ctypes.data, a.size,
                  cfg.r_particle, cfg.nb_dist)


def reconstruct_surface(cfg: SurfaceConfig, vertices, lib=None):
    """particles/mod.rs:220-274 from vertices: (DensityGrid, active blocks)."""
    lib = lib or _lib.load()
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    if v.shape[0] == 0:
        raise ValueError("point cloud has no vertices")
    bmin, span = bounding_box(v)
    if not (span > 0.0 and np.isfinite(span)):
        raise ValueError("degenerate point cloud (zero span)")
    placed = place(v, bmin, span, cfg, lib=lib)
    grid = DensityGrid(cfg.sx, cfg.sy, cfg.sz, capacity_blocks=max(placed.active.size, 1), lib=lib)
    k = grid.splat(placed.positions, cfg.r_scan())
    if k != placed.active.size:
        raise RuntimeError("splat activated %d blocks, placement %d" % (k, placed.active.size))
    finalize(grid, placed.active, cfg)
    return grid, placed.active


def reconstruct_and_smooth(cfg: SurfaceConfig, vertices, lib=None):
    """particles/mod.rs:279-302: reconstruct_surface, then n_smooth_iters
    mean-curvature sweeps clamped to [0, 1] on the Density slab."""
    grid, active = reconstruct_surface(cfg, vertices, lib=lib)
    if cfg.n_smooth_iters > 0 and active.size:
        a = np.ascontiguousarray(active, np.int64)
        grid.lib.call("msbg_sweep_density", grid.layout.h, grid.slab.ptr, a.ctypes.data, a.size, MEAN_CURVATURE,
                      cfg.smooth_dt, cfg.n_smooth_iters, 1)
    return grid, active
