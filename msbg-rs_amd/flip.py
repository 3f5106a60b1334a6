"""One FLIP-style step on the GPU over a MAC grid.

The reference has the pieces up to the divergence (fluid/{mod,mac,p2g,g2p}.rs)
and plans the pressure solve (docs/roadmap.md:610-702); this step composes
them with the band Poisson solve:

    Morton sort    msbg_morton_sort_particles (keys, permutation, sorted positions, velocities, masses)
    P2G            particles_to_grid (fluid/p2g.rs:141-164: the MAC sum-splat, P/M normalize)
    copy           the pre-projection faces u* (msbg_grid_copy)
    divergence     MacGrid::divergence (fluid/mac.rs:137-171) over the MAC value blocks
    solve          PoissonBand over the MAC blocks whose -x, -y, -z neighbours are MAC blocks
                   too (so every unknown cell has its six faces stored): A p = div to tol;
                   p = 0 outside the band (the air), zero gradient at the domain faces (walls)
    projection     u += grad p (msbg_mac_subtract_gradient with scale -1) over all MAC blocks,
                   so div u = div - A p, the solve's residual, on every band cell
    G2P            MacSampler on u and on u* and g2p_blend, per particle in one pass
                   (msbg_g2p_flip, fluid/g2p.rs:11-45)

Every stage runs in the HIP library; the step returns the new particle
velocities (in Morton order, with the sorted positions) and per-stage
HIP-event times.
"""
import numpy as np

from . import lib as _lib
from .device import DeviceGrid
from .solve import PoissonBand, Storage
from .transfer import MacGrid, g2p_flip, particles_to_grid


class FlipStep:
    """A MAC grid, its pre-projection copy and the pressure layout for one
    domain, reused across steps."""

    def __init__(self, sx, sy, sz, capacity_blocks, lib=None):
        self.lib = lib or _lib.load()
        self.sx, self.sy, self.sz = sx, sy, sz
        self.mac = MacGrid(sx, sy, sz, capacity_blocks, lib=self.lib)
        self.mac_tilde = MacGrid(sx, sy, sz, capacity_blocks, lib=self.lib)
        self.capacity_blocks = capacity_blocks

    def run(self, pos, vel, mass, r_p=2.0, alpha=0.95, storage=Storage.F32, refine=False, omega=1.9,
            sweeps_per_check=4, max_sweeps=20000, tol=1e-6, learned=None, phi=None, multigrid=False,
            compensated=False, direct=False):
        """One step.  pos / vel / mass: device arrays (n x 3, n x 3, n); the
        solve's arguments are PoissonBand.solve's.  compensated (refine or
        multigrid): the pressure is the compensated pair p + p_lo of two FP32
        fields through the solve and the projection, so the solve can reach a
        residual below one FP32 field's rounding floor.
        Returns (sorted positions, new velocities, stats) with stats holding
        the solve's msbg_poisson_stats, the active block count, the band's
        cell count and per-stage ms (HIP events)."""
        lib = self.lib
        n = pos.shape[0]
        ev = [lib.event() for _ in range(8)]
        rec = lambda i: lib.call("msbg_event_record", ev[i])
        rec(0)
        keys, perm = lib.empty(n, np.uint32), lib.empty(n, np.uint32)
        ps, vs, ms = lib.empty((n, 3), np.float32), lib.empty((n, 3), np.float32), lib.empty(n, np.float32)
        lib.call("msbg_morton_sort_particles", pos.ptr, vel.ptr, mass.ptr, n, keys.ptr, perm.ptr, ps.ptr, vs.ptr,
                 ms.ptr)
        rec(1)
        k = particles_to_grid(self.mac, ps, vs, ms, r_p)
        rec(2)
        for a in range(3):
            self.mac_tilde.velocity[a].copy_from(self.mac.velocity[a])
        active = self.mac.velocity[0].value_blocks()
        band_blocks = self.band_blocks(active)
        layout = DeviceGrid(self.sx, self.sy, self.sz, 4, 0.0, 0.0, capacity_blocks=max(active.size, 1), lib=lib)
        self.mac.divergence(layout, active)
        band = PoissonBand(layout, band_blocks, lib=lib)
        dev_slab, _ = layout.views()
        b = lib.empty(band.n_slots * 4096, np.float32)
        lib.call("msbg_memcpy_dtod", b.ptr, dev_slab, band.n_slots * 4096 * 4)
        rec(3)
        p = band.field(0)
        p_lo = band.field(0) if compensated else None
        st = band.solve(p, b, storage=storage, refine=refine, omega=omega, sweeps_per_check=sweeps_per_check,
                        max_sweeps=max_sweeps, tol=tol, learned=learned, phi=phi, multigrid=multigrid, x_lo=p_lo,
                        direct=direct)
        rec(4)
        self.mac.subtract_gradient(layout, p, active, -1.0, p_lo=p_lo)
        rec(5)
        out = g2p_flip(self.mac, self.mac_tilde, ps, vs, alpha)
        rec(6)
        lib.synchronize()
        names = ("sort", "p2g", "divergence", "solve", "projection", "g2p")
        stage_ms = {nm: lib.elapsed_ms(ev[i], ev[i + 1]) for i, nm in enumerate(names)}
        stage_ms["total"] = lib.elapsed_ms(ev[0], ev[6])
        for e in ev:
            lib.call("msbg_event_destroy", e)
        stats = {"active_blocks": int(k), "band_blocks": int(band_blocks.size), "cells": band.n_cells, "solve": st,
                 "ms": stage_ms}
        self.last = (layout, band, p, b, active, band_blocks)
        self.p_lo = p_lo
        self.sorted = (ps, vs, ms)
        return ps, out, stats

    def band_blocks(self, active):
        """The MAC blocks whose -x, -y and -z face neighbours are MAC blocks
        too (or lie outside the domain): the pressure unknowns."""
        nx, ny = (self.sx + 15) >> 4, (self.sy + 15) >> 4
        nz = (self.sz + 15) >> 4
        mac = np.zeros(nx * ny * nz, bool)
        mac[active] = True
        bx, by, bz = active % nx, (active // nx) % ny, active // (nx * ny)
        ok = ((bx == 0) | mac[np.maximum(active - 1, 0)]) & ((by == 0) | mac[np.maximum(active - nx, 0)]) & \
            ((bz == 0) | mac[np.maximum(active - nx * ny, 0)])
        return active[ok]

    def divergence_norms(self):
        """(||div u*||, ||div u||) over the band's cells after run(): the
        pre- and post-projection divergence (FP64, on the device)."""
        layout, band, p, b, active, _ = self.last
        post = DeviceGrid(self.sx, self.sy, self.sz, 4, 0.0, 0.0, capacity_blocks=max(active.size, 1), lib=self.lib)
        post.copy_from(layout)  # the band's slot table
        self.mac.divergence(post, active)
        d_slab, _ = post.views()
        zero = band.field(0)
        d = self.lib.empty(band.n_slots * 4096, np.float32)
        self.lib.call("msbg_memcpy_dtod", d.ptr, d_slab, band.n_slots * 4096 * 4)
        # ||d - A 0||^2 = ||d||^2 over the listed cells
        r2_post, b2_pre = band.residual(0, zero, d)[0], band.residual(0, zero, b)[0]
        return float(np.sqrt(b2_pre)), float(np.sqrt(r2_post))
