"""Host-side mirror of the reference's sparse grid (`src/sparse_grid.rs`).

`SparseGrid` keeps a host shadow of the block layout the HIP library works on:

* ``b2s[bid]`` (int32) names a block's slab slot: 0 is the shared empty dummy
  (an absent block, `BlockRef::Empty`), 1 the shared full dummy
  (`BlockRef::Full`), and ``>= 2`` an allocated value block;
* ``slab[slot]`` holds the block payload, ``B**3`` values with
  ``vid = vx | vy << L | vz << 2L`` (x fastest, `sparse_grid.rs:169-175`).

The method names and argument meaning follow `SparseGrid` (`sparse_grid.rs:107-435`);
the device copy (`DeviceGrid`) is made by `to_device()`.
"""

from __future__ import annotations

import numpy as np

EMPTY_SLOT = 0
FULL_SLOT = 1


class SparseGrid:
    """`SparseGrid<f32, BSX, N>` (sparse_grid.rs:61-155) with a numpy slab."""

    def __init__(self, name: str, sx: int, sy: int, sz: int, empty_value: float = 0.0,
                 full_value: float = 1.0, log2_bsx: int = 4, dtype=np.float32):
        self.name = name
        self.sx, self.sy, self.sz = int(sx), int(sy), int(sz)
        self.L = int(log2_bsx)
        self.B = 1 << self.L
        self.N = self.B ** 3
        # sparse_grid.rs:131-136
        self.nx = (self.sx + self.B - 1) >> self.L
        self.ny = (self.sy + self.B - 1) >> self.L
        self.nz = (self.sz + self.B - 1) >> self.L
        self.nxy = self.nx * self.ny
        self.n_blocks = self.nxy * self.nz
        self.dtype = np.dtype(dtype)
        self.b2s = np.zeros(self.n_blocks, dtype=np.int32)
        self._slab = np.zeros((8, self.N), dtype=self.dtype)
        self.n_slots = 2
        self._empty_value = self.dtype.type(empty_value)
        self._full_value = self.dtype.type(full_value)
        self._slab[EMPTY_SLOT] = self._empty_value
        self._slab[FULL_SLOT] = self._full_value

    # -- addressing (sparse_grid.rs:159-175, 217-220) ------------------------
    def get_block_id(self, x: int, y: int, z: int) -> int:
        return (x >> self.L) + (y >> self.L) * self.nx + (z >> self.L) * self.nxy

    def get_voxel_id(self, x: int, y: int, z: int) -> int:
        m = self.B - 1
        return (x & m) | ((y & m) << self.L) | ((z & m) << (2 * self.L))

    def get_block_coords_by_id(self, bid: int):
        return bid % self.nx, (bid // self.nx) % self.ny, bid // self.nxy

    # -- slab management -------------------------------------------------------
    @property
    def slab(self) -> np.ndarray:
        return self._slab[: self.n_slots]

    def _alloc_slots(self, k: int) -> np.ndarray:
        need = self.n_slots + k
        if need > self._slab.shape[0]:
            cap = max(need, 2 * self._slab.shape[0])
            grown = np.zeros((cap, self.N), dtype=self.dtype)
            grown[: self.n_slots] = self._slab[: self.n_slots]
            self._slab = grown
        slots = np.arange(self.n_slots, need, dtype=np.int32)
        self.n_slots = need
        return slots

    def empty_value(self):
        return self._empty_value

    def full_value(self):
        return self._full_value

    def set_empty_value(self, v: float) -> None:  # sparse_grid.rs:253-258
        self._empty_value = self.dtype.type(v)
        self._slab[EMPTY_SLOT] = self._empty_value

    def set_full_value(self, v: float) -> None:  # sparse_grid.rs:261-266
        self._full_value = self.dtype.type(v)
        self._slab[FULL_SLOT] = self._full_value

    def is_value_block(self, bid: int) -> bool:
        return self.b2s[bid] >= 2

    def is_full_block(self, bid: int) -> bool:
        return self.b2s[bid] == FULL_SLOT

    def is_empty_block(self, bid: int) -> bool:
        return self.b2s[bid] == EMPTY_SLOT

    def ensure_block(self, bid: int) -> None:  # sparse_grid.rs:314-324
        if self.b2s[bid] >= 2:
            return
        s = self._alloc_slots(1)[0]
        self._slab[s] = self._empty_value
        self.b2s[bid] = s

    def ensure_blocks_parallel(self, bids) -> None:  # sparse_grid.rs:333-355
        bids = np.asarray(bids, dtype=np.int64)
        new = bids[self.b2s[bids] < 2]
        if new.size == 0:
            return
        if np.unique(new).size != new.size:
            raise ValueError("ensure_blocks_parallel: bids must be unique")
        slots = self._alloc_slots(new.size)
        self._slab[slots] = self._empty_value
        self.b2s[new] = slots

    def fill_blocks_parallel(self, bids, value: float) -> None:  # sparse_grid.rs:359-369
        bids = np.asarray(bids, dtype=np.int64)
        s = self.b2s[bids]
        self._slab[s[s >= 2]] = self.dtype.type(value)

    def set_full_block(self, bid: int) -> None:  # sparse_grid.rs:278-282
        if self.b2s[bid] >= 2:
            raise ValueError("set_full_block on an allocated block")
        self.b2s[bid] = FULL_SLOT

    def set_empty_block(self, bid: int) -> None:  # sparse_grid.rs:270-274
        if self.b2s[bid] >= 2:
            raise ValueError("set_empty_block on an allocated block")
        self.b2s[bid] = EMPTY_SLOT

    # -- voxel access (sparse_grid.rs:286-400) ---------------------------------
    def get_voxel(self, x: int, y: int, z: int):
        return self._slab[self.b2s[self.get_block_id(x, y, z)], self.get_voxel_id(x, y, z)]

    def set_voxel(self, x: int, y: int, z: int, v: float) -> None:
        bid = self.get_block_id(x, y, z)
        if self.b2s[bid] < 2:  # allocate over None or a dummy (sparse_grid.rs:385-395)
            s = self._alloc_slots(1)[0]
            self._slab[s] = self._empty_value
            self.b2s[bid] = s
        self._slab[self.b2s[bid], self.get_voxel_id(x, y, z)] = self.dtype.type(v)

    def block_data(self, bid: int):
        """`get_block_data_mut`: a writable view of an allocated block, else None."""
        s = self.b2s[bid]
        return self._slab[s] if s >= 2 else None

    def value_blocks(self) -> np.ndarray:
        return np.nonzero(self.b2s >= 2)[0].astype(np.int64)

    # -- dense conversion helpers ----------------------------------------------
    def _dense_index(self):
        """(slot-major) index arrays mapping block voxels to dense (z, y, x)."""
        B, L = self.B, self.L
        v = np.arange(self.N)
        return v & (B - 1), (v >> L) & (B - 1), v >> (2 * L)

    def fill_from_function(self, bids, fn) -> None:
        """Allocate `bids` and set each voxel to fn(x, y, z) (vectorized numpy)."""
        bids = np.asarray(bids, dtype=np.int64)
        self.ensure_blocks_parallel(bids)
        vx, vy, vz = self._dense_index()
        bx, by, bz = bids % self.nx, (bids // self.nx) % self.ny, bids // self.nxy
        X = (bx[:, None] << self.L) + vx[None, :]
        Y = (by[:, None] << self.L) + vy[None, :]
        Z = (bz[:, None] << self.L) + vz[None, :]
        self._slab[self.b2s[bids]] = np.asarray(fn(X, Y, Z), dtype=self.dtype)

    def to_dense(self) -> np.ndarray:
        """Dense (sz, sy, sx) array of get_voxel over the whole domain."""
        x = np.arange(self.sx)
        y = np.arange(self.sy)
        z = np.arange(self.sz)
        Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
        bid = (X >> self.L) + (Y >> self.L) * self.nx + (Z >> self.L) * self.nxy
        m = self.B - 1
        vid = (X & m) | ((Y & m) << self.L) | ((Z & m) << (2 * self.L))
        return self._slab[self.b2s[bid], vid]

    def copy(self) -> "SparseGrid":
        g = SparseGrid(self.name, self.sx, self.sy, self.sz, float(self._empty_value),
                       float(self._full_value), self.L, self.dtype)
        g.b2s = self.b2s.copy()
        g._slab = self._slab[: self.n_slots].copy()
        g.n_slots = self.n_slots
        return g


def morton3(x: int, y: int, z: int) -> int:
    """`morton3` (multires/sort.rs:8-22): 21 bits per axis, x -> bit 0."""

    def spread(v: int) -> int:
        v &= 0x1FFFFF
        v = (v | (v << 32)) & 0x001F00000000FFFF
        v = (v | (v << 16)) & 0x001F0000FF0000FF
        v = (v | (v << 8)) & 0x100F00F00F00F00F
        v = (v | (v << 4)) & 0x10C30C30C30C30C3
        v = (v | (v << 2)) & 0x1249249249249249
        return v

    return spread(x) | (spread(y) << 1) | (spread(z) << 2)
