"""Particle-side entry points: Morton sort, bucketing, active blocks.

Mirrors `bucket_indices` / `BucketPlan` (particles/sort.rs:161-272),
`bucket_particles` (fluid/mod.rs:71-91), `active_blocks` (particles/active.rs:15,
fluid/p2g.rs:104) and `morton3` (multires/sort.rs:20) for particle voxels.
Inputs may be numpy arrays (uploaded) or DeviceArrays (used in place).
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import lib as _lib
from .lib import DeviceArray, load


def _dev(lib, a, dtype):
    if isinstance(a, DeviceArray):
        return a
    return lib.to_device(np.ascontiguousarray(a, dtype))


@dataclass
class BucketPlan:
    """particles/sort.rs:161-165 (device arrays; `n_occupied` occupied blocks)."""
    particle_blocks: DeviceArray
    starts: DeviceArray
    perm: DeviceArray
    n_occupied: int

    def host(self):
        k = self.n_occupied
        return self.particle_blocks.get()[:k], self.starts.get()[: k + 1], self.perm.get()


def morton_keys(pos, lib=None):
    lib = lib or load()
    p = _dev(lib, pos, np.float32)
    n = p.shape[0] if len(p.shape) > 1 else p.shape[0] // 3
    keys = lib.empty(n, np.uint32)
    lib.call("msbg_morton_keys", p.ptr, n, keys.ptr)
    return keys


def radix_sort_pairs(keys, vals, key_bits=32, lib=None):
    """Stable sort in place of device (keys, vals) by the low key_bits bits."""
    lib = lib or load()
    lib.call("msbg_radix_sort_pairs", keys.ptr, vals.ptr, keys.shape[0], key_bits)


def morton_sort(pos, lib=None):
    """(sorted 30-bit keys, permutation) as device arrays."""
    lib = lib or load()
    p = _dev(lib, pos, np.float32)
    n = p.shape[0]
    keys = lib.empty(n, np.uint32)
    perm = lib.empty(n, np.uint32)
    lib.call("msbg_morton_sort", p.ptr, n, keys.ptr, perm.ptr)
    return keys, perm


def morton_sort_particles(pos, vel, mass=None, lib=None):
    """Morton-sorted particle set: (keys, perm, positions, velocities,
    masses) as device arrays, the arrays in sorted order (one packed record
    gathered per particle; the same values as morton_sort plus gather_rows)."""
    lib = lib or load()
    p, v = _dev(lib, pos, np.float32), _dev(lib, vel, np.float32)
    m = _dev(lib, mass, np.float32) if mass is not None else None
    n = p.shape[0]
    keys, perm = lib.empty(n, np.uint32), lib.empty(n, np.uint32)
    ps, vs, ms = lib.empty((n, 3), np.float32), lib.empty((n, 3), np.float32), lib.empty(n, np.float32)
    lib.call("msbg_morton_sort_particles", p.ptr, v.ptr, m.ptr if m is not None else None, n, keys.ptr, perm.ptr,
             ps.ptr, vs.ptr, ms.ptr)
    return keys, perm, ps, vs, ms


def particle_bids(sx, sy, sz, pos, lib=None):
    lib = lib or load()
    p = _dev(lib, pos, np.float32)
    n = p.shape[0]
    bids = lib.empty(n, np.int32)
    lib.call("msbg_particle_bids", sx, sy, sz, p.ptr, n, bids.ptr)
    return bids


def bucket_indices(bids, n_blocks, lib=None) -> BucketPlan:
    """particles/sort.rs:174: stable counting order by block id."""
    lib = lib or load()
    b = _dev(lib, bids, np.int32)
    n = b.shape[0]
    perm = lib.empty(max(n, 1), np.int32)
    blocks = lib.empty(max(n, 1), np.int32)
    starts = lib.empty(n + 1, np.int32)
    k = C.c_int64()
    lib.call("msbg_bucket_indices", b.ptr, n, n_blocks, perm.ptr, blocks.ptr, starts.ptr, C.byref(k))
    if n == 0:
        perm = lib.empty(0, np.int32)
    return BucketPlan(blocks, starts, perm, k.value)


def active_blocks(sx, sy, sz, pos, r_scan, lib=None) -> np.ndarray:
    """particles/active.rs:15-47: sorted union of footprint blocks (host int64)."""
    lib = lib or load()
    p = _dev(lib, pos, np.float32)
    n = p.shape[0]
    nb = ((sx + 15) >> 4) * ((sy + 15) >> 4) * ((sz + 15) >> 4)
    out = lib.empty(nb, np.int32)
    k = C.c_int64()
    lib.call("msbg_active_blocks", sx, sy, sz, p.ptr, n, r_scan, out.ptr, C.byref(k))
    return out.get()[: k.value].astype(np.int64)


def gather_rows(src: DeviceArray, perm: DeviceArray, lib=None) -> DeviceArray:
    """dst[i] = src[perm[i]] row-wise (fluid/mod.rs:84-89)."""
    lib = lib or load()
    n = perm.shape[0]
    row_bytes = src.nbytes // max(src.shape[0], 1)
    dst = lib.empty((n,) + src.shape[1:], src.dtype)
    lib.call("msbg_gather_rows", src.ptr, perm.ptr, n, row_bytes, dst.ptr)
    return dst


class DensityGrid:
    """A Density (u16) channel over a 16^3-block layout grid: the layout's
    bid -> slot table and a device u16 slab with a slot for every block the
    layout can hold (SparseGrid<Density, 16, 4096>, empty Density(0), full
    Density(u16::MAX))."""

    def __init__(self, sx, sy, sz, capacity_blocks, lib=None):
        from .device import DeviceGrid
        self.lib = lib or _lib.load()
        self.layout = DeviceGrid(sx, sy, sz, 4, 0.0, 1.0, capacity_blocks=capacity_blocks, lib=self.lib)
        self.slots = capacity_blocks + 2
        self.slab = self.lib.empty((self.slots, 4096), np.uint16)

    def splat(self, pos, r_scan):
        """particles::splat (particles/splat.rs:92-319): returns the number of
        active blocks."""
        d = pos if isinstance(pos, DeviceArray) else self.lib.to_device(np.ascontiguousarray(pos, np.float32))
        k = C.c_int64()
        self.lib.call("msbg_density_splat", self.layout.h, self.slab.ptr, self.slots, d.ptr, d.shape[0], r_scan,
                      C.byref(k))
        return k.value

    def to_host(self):
        """(b2s, u16 slab[n_slots])"""
        g = self.layout.to_host()
        return g.b2s, self.slab.get()[: g.n_slots]
