"""Device-side mirror of the reference's grid and sweeper entry points.

`DeviceGrid` is a `SparseGrid` resident in HBM (the C-ABI `msbg_grid`),
`Sweeper` / `apply_pde` / `PdeParams` / `Stencil` mirror `src/solver.rs:43-240`,
`halo_fill` mirrors `HaloBlock::fill` (multires/halo.rs:31) batched over
blocks, and the multires helpers mirror `fill_multires`,
`downsample_channel_avg` and `regularize_refinement_map`.
"""
import ctypes as C
import enum
from dataclasses import dataclass

import numpy as np

from .grid import SparseGrid
from .lib import DeviceArray, load, ptr


class Stencil(enum.IntEnum):
    """solver.rs:43-51."""
    Laplacian = 0
    MeanCurvature = 1
    BiLaplacian = 2


class BoundaryCondition(enum.IntEnum):
    """math/boundary.rs:6-14."""
    Clamp = 0
    Neumann = 1
    Dirichlet = 2


@dataclass
class PdeParams:
    """solver.rs:56-61."""
    dt: float
    iterations: int
    do_constr_zero_one: bool = False


def _i64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int64).ravel())


class DeviceGrid:
    """A device-resident `SparseGrid<f32>` / multires level (C-ABI msbg_grid)."""

    def __init__(self, sx, sy, sz, log2_bsx=4, empty_value=0.0, full_value=1.0, capacity_blocks=1024, lib=None):
        self.lib = lib or load()
        h = C.c_void_p()
        self.lib.call("msbg_grid_new", sx, sy, sz, log2_bsx, empty_value, full_value, capacity_blocks, C.byref(h))
        self.h = h.value
        self.L = log2_bsx
        self.B = 1 << log2_bsx
        self.N = self.B ** 3
        self.sx, self.sy, self.sz = sx, sy, sz
        d = self.dims()
        self.nx, self.ny, self.nz, self.n_blocks = int(d[3]), int(d[4]), int(d[5]), int(d[6])
        self.nxy = self.nx * self.ny

    @classmethod
    def from_host(cls, host: SparseGrid, capacity_blocks=None, lib=None):
        cap = capacity_blocks if capacity_blocks is not None else max(host.n_slots - 2, 1)
        g = cls(host.sx, host.sy, host.sz, host.L, float(host.empty_value()), float(host.full_value()), cap, lib)
        slab = np.ascontiguousarray(host.slab, np.float32)
        g.lib.call("msbg_grid_load", g.h, host.b2s.ctypes.data, slab.ctypes.data, host.n_slots)
        return g

    def to_host(self, name="grid") -> SparseGrid:
        d = self.dims()
        n_slots = int(d[7])
        host = SparseGrid(name, self.sx, self.sy, self.sz, 0.0, 1.0, self.L)
        b2s = np.zeros(self.n_blocks, np.int32)
        slab = np.zeros((n_slots, self.N), np.float32)
        ns = C.c_int64()
        self.lib.call("msbg_grid_store", self.h, b2s.ctypes.data, slab.ctypes.data, n_slots, C.byref(ns))
        host.b2s = b2s
        host._slab = slab
        host.n_slots = n_slots
        host._empty_value = slab[0, 0]
        host._full_value = slab[1, 0]
        return host

    def copy_from(self, src: "DeviceGrid"):
        """Deep copy of src's layout and values (msbg_grid_copy)."""
        self.lib.call("msbg_grid_copy", self.h, src.h)

    def block_slots(self):
        """Host copy of the bid -> slot table (no device traffic)."""
        b2s = np.zeros(self.n_blocks, np.int32)
        self.lib.call("msbg_grid_block_slots", self.h, b2s.ctypes.data)
        return b2s

    def value_blocks(self):
        """Sorted bids of the value blocks (slot >= 2)."""
        return np.nonzero(self.block_slots() >= 2)[0].astype(np.int64)

    def dims(self):
        d = np.zeros(8, np.int64)
        self.lib.call("msbg_grid_dims", self.h, d.ctypes.data)
        return d

    def views(self):
        s, b = C.c_void_p(), C.c_void_p()
        self.lib.call("msbg_grid_device_views", self.h, C.byref(s), C.byref(b))
        return s.value, b.value

    # sparse_grid.rs entry points
    def ensure_blocks_parallel(self, bids):
        b = _i64(bids)
        self.lib.call("msbg_grid_ensure_blocks", self.h, ptr(b), b.size)

    def fill_blocks_parallel(self, bids, value):
        b = _i64(bids)
        self.lib.call("msbg_grid_fill_blocks", self.h, ptr(b), b.size, value)

    def set_full_block(self, bid):
        self.lib.call("msbg_grid_set_full_block", self.h, int(bid))

    def set_empty_block(self, bid):
        self.lib.call("msbg_grid_set_empty_block", self.h, int(bid))

    def set_empty_value(self, v):
        self.lib.call("msbg_grid_set_empty_value", self.h, v)

    def set_full_value(self, v):
        self.lib.call("msbg_grid_set_full_value", self.h, v)

    def write_blocks(self, bids, data):
        b = _i64(bids)
        d = np.ascontiguousarray(data, np.float32)
        assert d.size == b.size * self.N
        self.lib.call("msbg_grid_write_blocks", self.h, ptr(b), b.size, ptr(d))

    def read_blocks(self, bids):
        b = _i64(bids)
        out = np.zeros((b.size, self.N), np.float32)
        self.lib.call("msbg_grid_read_blocks", self.h, ptr(b), b.size, ptr(out))
        return out

    def value_block_ptr(self, bid):
        p = C.c_void_p()
        self.lib.call("msbg_grid_value_block_ptr", self.h, int(bid), C.byref(p))
        return p.value

    def free(self):
        if getattr(self, "h", None):
            self.lib.call("msbg_grid_free", self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def color_buckets(grid: DeviceGrid, active):
    """build_color_buckets (solver.rs:247-267): (concatenated order, counts[8])."""
    a = _i64(active)
    out = np.zeros(max(a.size, 1), np.int64)
    counts = np.zeros(8, np.int64)
    grid.lib.call("msbg_color_buckets", grid.h, ptr(a), a.size, ptr(out), ptr(counts))
    return out[: a.size], counts


class Sweeper:
    """`Sweeper<f32, 16, 4096, HSX>` (solver.rs:141-225) over a DeviceGrid.

    `num_threads` and `fence` of the reference constructor have no meaning on
    the device and are accepted for signature parity only.
    """

    def __init__(self, grid: DeviceGrid, num_threads=None, fence=None):
        self.grid = grid

    def sweep(self, active, stencil: Stencil, params: PdeParams):
        a = _i64(active)
        self.grid.lib.call("msbg_sweep", self.grid.h, ptr(a), a.size, int(stencil), params.dt, params.iterations,
                           int(bool(params.do_constr_zero_one)))


def apply_pde(grid: DeviceGrid, active, stencil: Stencil, params: PdeParams, num_threads=None, fence=None):
    """solver.rs:228-240."""
    Sweeper(grid, num_threads, fence).sweep(active, stencil, params)


class SweepPlan:
    """Colour buckets resident on the device, re-run without host work.

    Single level (`coarse` None), fine blocks with cross-level halos
    (`coarse` = level 1, `levels` = the device level map), or every level of a
    multires grid (`level_grids` = [level 0, level 1, ...], `levels`): the
    active blocks are then swept at their own levels.
    """

    def __init__(self, grid: DeviceGrid, active, coarse: DeviceGrid = None, levels: DeviceArray = None,
                 level_grids=None):
        self.grid, self.coarse, self.levels = grid, coarse, levels
        a = _i64(active)
        h = C.c_void_p()
        if level_grids is not None:
            self._grids = list(level_grids)
            arr = (C.c_void_p * len(self._grids))(*[x.h for x in self._grids])
            grid.lib.call("msbg_sweep_plan_new_levels", arr, len(self._grids), levels.ptr, ptr(a), a.size,
                          C.byref(h))
        elif coarse is None:
            grid.lib.call("msbg_sweep_plan_new", grid.h, ptr(a), a.size, C.byref(h))
        else:
            grid.lib.call("msbg_sweep_plan_new_multires", grid.h, coarse.h, levels.ptr, ptr(a), a.size, C.byref(h))
        self.h = h.value

    def run_async(self, stencil, dt, iterations, clamp01=False):
        self.grid.lib.call("msbg_sweep_plan_run_async", self.h, int(stencil), dt, iterations, int(clamp01))

    def run_graph_async(self, stencil, dt, iterations, clamp01=False):
        """The same passes replayed as one cached HIP graph."""
        self.grid.lib.call("msbg_sweep_plan_run_graph_async", self.h, int(stencil), dt, iterations, int(clamp01))

    def run_timed(self, stencil, dt, iterations, clamp01=False):
        """Runs and returns (summed kernel ms, launch count) from per-launch HIP events."""
        ms, n = C.c_float(), C.c_int64()
        self.grid.lib.call("msbg_sweep_plan_run_timed", self.h, int(stencil), dt, iterations, int(clamp01),
                           C.byref(ms), C.byref(n))
        return ms.value, n.value

    def free(self):
        if getattr(self, "h", None):
            self.grid.lib.call("msbg_sweep_plan_free", self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def halo_fill(grid: DeviceGrid, bids, halo: int, full: bool, bc=BoundaryCondition.Neumann):
    """HaloBlock::fill::<halo, full> for each bid: (n, HSX, HSX, HSX) tiles, [z, y, x]."""
    b = _i64(bids)
    hsx = 16 + 2 * halo
    tiles = grid.lib.empty((b.size, hsx, hsx, hsx), np.float32)
    grid.lib.call("msbg_halo_gather", grid.h, ptr(b), b.size, halo, int(full), int(bc), tiles.ptr)
    return tiles.get()


def halo_topology(grid: DeviceGrid, bids, halo: int, coarse: DeviceGrid = None, levels: DeviceArray = None):
    """Neighbour slot codes per bid, (n, (2H+1)^3) int32 (see msbg_halo_topology)."""
    b = _i64(bids)
    nb = (2 * halo + 1) ** 3
    out = np.zeros((b.size, nb), np.int32)
    grid.lib.call("msbg_halo_topology", grid.h, coarse.h if coarse else None, levels.ptr if levels else None, ptr(b),
                  b.size, halo, ptr(out))
    return out


def halo_fill_multires(fine: DeviceGrid, coarse: DeviceGrid, levels: DeviceArray, bids, halo, full,
                       bc=BoundaryCondition.Neumann):
    """fill_multires (multires/halo.rs:187) for each bid."""
    b = _i64(bids)
    hsx = 16 + 2 * halo
    tiles = fine.lib.empty((b.size, hsx, hsx, hsx), np.float32)
    fine.lib.call("msbg_halo_gather_multires", fine.h, coarse.h, levels.ptr, ptr(b), b.size, halo, int(full),
                  int(bc), tiles.ptr)
    return tiles.get()


def sweep_multires(fine: DeviceGrid, coarse: DeviceGrid, levels: DeviceArray, active, stencil, params: PdeParams):
    a = _i64(active)
    fine.lib.call("msbg_sweep_multires", fine.h, coarse.h, levels.ptr, ptr(a), a.size, int(stencil), params.dt,
                  params.iterations, int(bool(params.do_constr_zero_one)))


def sweep_levels(level_grids, levels: DeviceArray, active, stencil, params: PdeParams):
    """The 8-colour sweep of every active block at its own level (level l of
    `level_grids` has block edge 16 >> l); coarse-level halos read finer
    neighbours as 2^3 means and coarser ones by injection."""
    a = _i64(active)
    arr = (C.c_void_p * len(level_grids))(*[x.h for x in level_grids])
    level_grids[0].lib.call("msbg_sweep_levels", arr, len(level_grids), levels.ptr, ptr(a), a.size, int(stencil),
                            params.dt, params.iterations, int(bool(params.do_constr_zero_one)))


def downsample_channel_avg(coarse: DeviceGrid, fine: DeviceGrid, bids):
    """multires/downsample.rs:13-58 for `bids`."""
    b = _i64(bids)
    fine.lib.call("msbg_downsample_avg", coarse.h, fine.h, ptr(b), b.size)


def regularize_refinement_map(levels: DeviceArray, nx, ny, nz, n_levels):
    """multires/refinement.rs:66-107, in place on a device level map."""
    levels.lib.call("msbg_regularize_levels", levels.ptr, nx, ny, nz, n_levels)


class Topology:
    """multires/refinement.rs:110-116: FINE_COARSE blocks per level (host
    int64 arrays, ascending) and the EXISTS block count."""

    def __init__(self, blocks_fine_coarse, n_act_blocks):
        self.blocks_fine_coarse = blocks_fine_coarse
        self.n_act_blocks = n_act_blocks


def compute_block_topology(levels: DeviceArray, nx, ny, nz, n_levels, block_flags: DeviceArray = None):
    """compute_block_topology (multires/refinement.rs:124-208) on a device
    level map: (device u16 block flags -- the reference's store.flags[0] --,
    Topology).  store.level0 is the level map itself."""
    lib = levels.lib
    nb = nx * ny * nz
    flags = lib.empty(nb, np.uint16)
    fc = lib.empty(nb, np.int64)
    counts = np.zeros(n_levels, np.int64)
    n_act = C.c_int64()
    lib.call("msbg_block_topology", levels.ptr, nx, ny, nz, n_levels,
             None if block_flags is None else block_flags.ptr, flags.ptr, fc.ptr, counts.ctypes.data,
             C.byref(n_act))
    allfc = fc.get()
    lists, k = [], 0
    for c in counts:
        lists.append(allfc[k:k + int(c)].copy())
        k += int(c)
    return flags, Topology(lists, n_act.value)
