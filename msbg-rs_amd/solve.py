"""Poisson relaxation on a band of value blocks (host side of the
msbg_poisson_* entry points).

There is no reference counterpart: the reference plans the solver
(docs/roadmap.md:610-702, weighted Jacobi at :622-626) but does not build it.
Fields are device slabs shaped like the layout grid's slab (n_slots x 4096
elements, slot-major) in one of three storage types; arithmetic is FP32.
"""
import ctypes as C

import numpy as np

from . import lib as _lib


class Storage:
    F32, F16, BF16 = 0, 1, 2


# numpy dtype holding each storage type's elements (bf16 as raw uint16 bits)
STORAGE_DTYPE = {Storage.F32: np.float32, Storage.F16: np.float16, Storage.BF16: np.uint16}


def to_storage(a, storage):
    """float32 array -> numpy array of the storage type (RNE), on the host."""
    a = np.ascontiguousarray(a, np.float32)
    if storage == Storage.F32:
        return a.copy()
    if storage == Storage.F16:
        return a.astype(np.float16)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    r[nan] = ((u[nan] >> 16) | 0x40).astype(np.uint16)
    return r


def from_storage(a, storage):
    """numpy array of the storage type -> float32."""
    if storage == Storage.F32:
        return np.asarray(a, np.float32)
    if storage == Storage.F16:
        return np.asarray(a, np.float16).astype(np.float32)
    return (np.asarray(a, np.uint16).astype(np.uint32) << 16).view(np.float32)


class PoissonBand:
    """The 7-point Poisson operator over the listed value blocks of a 16^3-block
    layout grid (a DeviceGrid).  Absent / full neighbour blocks read slot 0 / 1
    of the field, the domain boundary mirrors (Neumann)."""

    def __init__(self, layout, active, lib=None):
        self.lib = lib or layout.lib
        self.active = np.ascontiguousarray(active, np.int64)
        h = C.c_void_p()
        self.lib.call("msbg_poisson_new", layout.h, self.active.ctypes.data, self.active.size, C.byref(h))
        self.h = h.value
        a, c, s = C.c_int64(), C.c_int64(), C.c_int64()
        self.lib.call("msbg_poisson_info", self.h, C.byref(a), C.byref(c), C.byref(s))
        self.n_active, self.n_cells, self.n_slots = a.value, c.value, s.value

    def field(self, storage=Storage.F32, host=None):
        """A device field (DeviceArray of n_slots * 4096 elements), zero or from `host`."""
        n = self.n_slots * 4096
        if host is None:
            host = np.zeros(n, STORAGE_DTYPE[storage])
        host = np.ascontiguousarray(host).reshape(-1)
        assert host.size == n and host.dtype == np.dtype(STORAGE_DTYPE[storage])
        return self.lib.to_device(host)

    def jacobi(self, storage, x_in, x_out, b, omega):
        self.lib.call("msbg_poisson_jacobi", self.h, storage, x_in.ptr, x_out.ptr, b.ptr, omega)

    def rbgs(self, storage, x, b, color, omega):
        self.lib.call("msbg_poisson_rbgs", self.h, storage, x.ptr, b.ptr, color, omega)

    def residual(self, storage, x, b, r=None):
        """(||b - A x||^2, ||b||^2) in FP64; r (FP32 field) receives b - A x."""
        r2, b2 = C.c_double(), C.c_double()
        self.lib.call("msbg_poisson_residual", self.h, storage, x.ptr, b.ptr, None if r is None else r.ptr,
                      C.byref(r2), C.byref(b2))
        return r2.value, b2.value

    def solve(self, x, b, storage=Storage.F32, refine=False, omega=1.9, sweeps_per_check=10, max_sweeps=10000,
              tol=1e-6):
        """Red-black SOR solve of A x = b (x, b FP32 fields; x = initial guess,
        updated in place).  refine=True: FP32 iterative refinement with the
        inner sweeps in `storage`.  Returns the msbg_poisson_stats fields."""
        prm = _lib.PoissonParams(storage, int(refine), omega, sweeps_per_check, max_sweeps, 0, tol)
        st = _lib.PoissonStats()
        self.lib.call("msbg_poisson_solve", self.h, C.byref(prm), x.ptr, b.ptr, C.byref(st))
        return {k: getattr(st, k) for k, _ in _lib.PoissonStats._fields_ if k != "reserved"}

    def close(self):
        if self.h:
            self.lib.call("msbg_poisson_free", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
