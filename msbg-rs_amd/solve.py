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
        """A device field (DeviceArray of n_slots * 4096 elements), zero (filled
        on the device) or from `host`."""
        n = self.n_slots * 4096
        if host is None:
            f = self.lib.empty(n, STORAGE_DTYPE[storage])
            self.lib.call("msbg_fill_u32", f.ptr, f.nbytes // 4, 0)
            return f
        host = np.ascontiguousarray(host).reshape(-1)
        assert host.size == n and host.dtype == np.dtype(STORAGE_DTYPE[storage])
        return self.lib.to_device(host)

    def jacobi(self, storage, x_in, x_out, b, omega):
        self.lib.call("msbg_poisson_jacobi", self.h, storage, x_in.ptr, x_out.ptr, b.ptr, omega)

    def jacobi_sweeps(self, storage, x, tmp, b, omega, sweeps):
        """`sweeps` (even) Jacobi sweeps x -> tmp -> x as one HIP graph; tmp
        must hold x's values outside the listed cells."""
        self.lib.call("msbg_poisson_jacobi_sweeps", self.h, storage, x.ptr, tmp.ptr, b.ptr, omega, sweeps)

    def rbgs(self, storage, x, b, color, omega):
        self.lib.call("msbg_poisson_rbgs", self.h, storage, x.ptr, b.ptr, color, omega)

    def residual(self, storage, x, b, r=None):
        """(||b - A x||^2, ||b||^2) in FP64; r (FP32 field) receives b - A x."""
        r2, b2 = C.c_double(), C.c_double()
        self.lib.call("msbg_poisson_residual", self.h, storage, x.ptr, b.ptr, None if r is None else r.ptr,
                      C.byref(r2), C.byref(b2))
        return r2.value, b2.value

    def residual_pair(self, x, x_lo, b, r=None):
        """(||b - A (x + x_lo)||^2, ||b||^2) of a compensated FP32 iterate; r
        (FP32 field) receives the residual, rounded once."""
        r2, b2 = C.c_double(), C.c_double()
        self.lib.call("msbg_poisson_residual_pair", self.h, x.ptr, x_lo.ptr, b.ptr, None if r is None else r.ptr,
                      C.byref(r2), C.byref(b2))
        return r2.value, b2.value

    def solve(self, x, b, storage=Storage.F32, refine=False, omega=1.9, sweeps_per_check=10, max_sweeps=10000,
              tol=1e-6, learned=None, phi=None, multigrid=False, x_lo=None, direct=False):
        """Red-black SOR solve of A x = b (x, b FP32 fields; x = initial guess,
        updated in place).  refine=True: FP32 iterative refinement with the
        inner sweeps in `storage`.  multigrid=True: FP32 iterative refinement
        whose inner solve is one multigrid V-cycle per check, with
        sweeps_per_check smoothing sweeps (factor omega; 1.0 = Gauss-Seidel)
        before and after the coarse correction on every level, the band's in
        `storage`; max_sweeps caps the band's sweeps (2 sweeps_per_check per
        cycle).  learned (a CorrectorNet) with phi (the phase-field
        DeviceGrid): a correction is tried after every check period.  Returns
        the msbg_poisson_stats fields.  x_lo (an FP32 field, refine or multigrid
        only): the low part of a compensated iterate x + x_lo, carried through
        the refinement with error-free FP32 sums (x_lo zero on entry for a
        fresh solve).  direct: every check period's kernels are launched
        directly instead of replayed as a captured graph (same results)."""
        mode = 2 if multigrid else int(refine)
        prm = _lib.PoissonParams(storage, mode, omega, sweeps_per_check, max_sweeps, 1 if direct else 0, tol,
                                 learned.h if learned is not None else None, phi.h if phi is not None else None,
                                 x_lo.ptr if x_lo is not None else None)
        st = _lib.PoissonStats()
        self.lib.call("msbg_poisson_solve", self.h, C.byref(prm), x.ptr, b.ptr, C.byref(st))
        return {k: getattr(st, k) for k, _ in _lib.PoissonStats._fields_}

    def net_input(self, phi, r, inv_scale):
        """Network input of every listed block (device bf16 bits, n x 20^3 x 2)."""
        x0 = self.lib.empty((self.n_active, 20, 20, 20, 2), np.uint16)
        self.lib.call("msbg_poisson_net_input", self.h, phi.h, r.ptr, inv_scale, x0.ptr)
        return x0

    def error_target(self, x, xstar, inv_scale, out=None):
        """(x* - x) * inv_scale on every listed block (device float32, n x 4096)."""
        t = out if out is not None else self.lib.empty((self.n_active, 4096), np.float32)
        self.lib.call("msbg_poisson_error_target", self.h, x.ptr, xstar.ptr, inv_scale, t.ptr)
        return t

    def learned_correct(self, net, phi, x, b):
        """One learned correction of x; returns (r2_before, r2_after, accepted)."""
        r2b, r2a, acc = C.c_double(), C.c_double(), C.c_int()
        self.lib.call("msbg_poisson_learned_correct", self.h, net.h, phi.h, x.ptr, b.ptr, C.byref(r2b), C.byref(r2a),
                      C.byref(acc))
        return r2b.value, r2a.value, bool(acc.value)

    def close(self):
        if self.h:
            self.lib.call("msbg_poisson_free", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
