"""Grid <-> particle transfers on the device.

`MacGrid`, `particles_to_grid`, `MacSampler`, `g2p_blend` mirror
fluid/mac.rs:31-134, fluid/p2g.rs:141-164 and fluid/g2p.rs:11-45; `sample`
mirrors Sampler::sample (math/sample.rs:627-658).  `TransferGrid` carries the
trilinear / quadratic B-spline P2G and G2P (no reference counterpart).
"""
import ctypes as C
import enum

import numpy as np

from .device import BoundaryCondition, DeviceGrid
from .lib import DeviceArray, load


class Interpolation(enum.IntEnum):
    """math/boundary.rs:25-30."""
    Linear = 0
    CubicBSpline = 1


class GridAlignment(enum.IntEnum):
    """math/boundary.rs:16-23."""
    Corner = 0
    CellCentered = 1


def _dev(lib, a, dtype=np.float32):
    if isinstance(a, DeviceArray):
        return a
    return lib.to_device(np.ascontiguousarray(a, dtype))


def _handles(grids):
    arr = (C.c_void_p * len(grids))(*[g.h for g in grids])
    return arr


def sample(grid: DeviceGrid, pos, interp=Interpolation.Linear, align=GridAlignment.Corner,
           bc=BoundaryCondition.Clamp) -> DeviceArray:
    """Sampler::sample for every row of pos (n x 3): device array of n values."""
    lib = grid.lib
    p = _dev(lib, pos)
    n = p.shape[0]
    out = lib.empty(n, np.float32)
    lib.call("msbg_sample", grid.h, p.ptr, n, int(interp), int(align), int(bc), out.ptr)
    return out


class MacGrid:
    """fluid/mac.rs:31-46: three face-velocity and three face-mass channels."""

    def __init__(self, sx, sy, sz, capacity_blocks, lib=None):
        self.lib = lib or load()
        mk = lambda: DeviceGrid(sx, sy, sz, 4, 0.0, 0.0, capacity_blocks, lib=self.lib)
        self.velocity = [mk(), mk(), mk()]
        self.mass = [mk(), mk(), mk()]
        self.sx, self.sy, self.sz = sx, sy, sz

    def divergence(self, dst: DeviceGrid, active):
        """MacGrid::divergence (fluid/mac.rs:137-171) into dst over `active` bids."""
        a = np.ascontiguousarray(active, np.int64)
        self.lib.call("msbg_mac_divergence", _handles(self.velocity), dst.h, a.ctypes.data, a.size)

    def subtract_gradient(self, p_layout: DeviceGrid, p, active, scale=1.0, p_lo=None):
        """Face velocities of the `active` blocks minus scale * grad p (p a
        field over p_layout's slots, e.g. a PoissonBand solution; p_lo: the low
        part of a compensated solution p + p_lo)."""
        a = np.ascontiguousarray(active, np.int64)
        if p_lo is None:
            self.lib.call("msbg_mac_subtract_gradient", _handles(self.velocity), p_layout.h, p.ptr, a.ctypes.data,
                          a.size, scale)
        else:
            self.lib.call("msbg_mac_subtract_gradient_pair", _handles(self.velocity), p_layout.h, p.ptr, p_lo.ptr,
                          a.ctypes.data, a.size, scale)


def particles_to_grid(mac: MacGrid, pos, vel, mass, r_p) -> int:
    """fluid/p2g.rs:141-164; returns the number of active blocks."""
    lib = mac.lib
    p, v, m = _dev(lib, pos), _dev(lib, vel), _dev(lib, mass)
    k = C.c_int64()
    lib.call("msbg_particles_to_grid_mac", _handles(mac.velocity), _handles(mac.mass), p.ptr, v.ptr, m.ptr,
             p.shape[0], r_p, C.byref(k))
    return k.value


class MacSampler:
    """fluid/g2p.rs:11-33."""

    def __init__(self, mac: MacGrid, bc=BoundaryCondition.Clamp):
        self.mac, self.bc = mac, bc

    def sample(self, pos, interp=Interpolation.Linear) -> DeviceArray:
        lib = self.mac.lib
        p = _dev(lib, pos)
        out = lib.empty((p.shape[0], 3), np.float32)
        lib.call("msbg_mac_sample", _handles(self.mac.velocity), p.ptr, p.shape[0], int(interp), int(self.bc), out.ptr)
        return out


def g2p_blend(u_old, u, u_tilde, alpha, lib=None) -> DeviceArray:
    """fluid/g2p.rs:42-45."""
    lib = lib or load()
    a, b, c = _dev(lib, u_old), _dev(lib, u), _dev(lib, u_tilde)
    out = lib.empty(a.shape, np.float32)
    lib.call("msbg_g2p_blend", a.ptr, b.ptr, c.ptr, a.shape[0], alpha, out.ptr)
    return out


def g2p_flip(mac: MacGrid, mac_tilde: MacGrid, pos, u_old, alpha, interp=Interpolation.Linear,
             bc=BoundaryCondition.Clamp) -> DeviceArray:
    """fluid/g2p.rs:11-45 per particle: the samples of u (mac) and u*
    (mac_tilde) and the FLIP/PIC blend, in one pass (the bits of
    MacSampler.sample twice and g2p_blend)."""
    lib = mac.lib
    p, uo = _dev(lib, pos), _dev(lib, u_old)
    out = lib.empty((p.shape[0], 3), np.float32)
    lib.call("msbg_g2p_flip", _handles(mac.velocity), _handles(mac_tilde.velocity), p.ptr, uo.ptr, p.shape[0],
             int(interp), int(bc), alpha, out.ptr)
    return out


class TransferGrid:
    """Cell-centred mass + velocity channels for the B-spline P2G / G2P."""

    def __init__(self, sx, sy, sz, capacity_blocks, lib=None):
        self.lib = lib or load()
        mk = lambda: DeviceGrid(sx, sy, sz, 4, 0.0, 0.0, capacity_blocks, lib=self.lib)
        self.mass = mk()
        self.velocity = [mk(), mk(), mk()]

    def p2g(self, order, pos, vel, mass=None) -> int:
        lib = self.lib
        p, v = _dev(lib, pos), _dev(lib, vel)
        m = _dev(lib, mass) if mass is not None else None
        k = C.c_int64()
        lib.call("msbg_p2g_bspline", order, self.mass.h, _handles(self.velocity), p.ptr, v.ptr,
                 m.ptr if m is not None else None, p.shape[0], C.byref(k))
        return k.value

    def g2p(self, order, pos) -> DeviceArray:
        lib = self.lib
        p = _dev(lib, pos)
        out = lib.empty((p.shape[0], 3), np.float32)
        lib.call("msbg_g2p_bspline", order, _handles(self.velocity), p.ptr, p.shape[0], out.ptr)
        return out
