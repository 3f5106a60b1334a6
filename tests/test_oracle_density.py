"""Pins of the oracle's Density (u16) sweep: the reference's test_h_11
dequant known answer (multires/halo.rs:627-645), constant-field fixed points,
and agreement with the FP32 sweep to the quantisation step."""
import numpy as np
import pytest

from test_gpu_sweep import random_grid


def density_of(g):
    """u16 slab of a [0, 1] FP32 grid: round(v * 65535) (Density::from_f32)."""
    return np.floor(np.clip(np.asarray(g.slab, np.float64), 0, 1) * 65535.0 + 0.5).astype(np.uint16)


def sweep16(orc, g, slab16, active, stencil, dt, it=1, clamp01=1):
    out = np.ascontiguousarray(slab16).copy()
    act = np.ascontiguousarray(active, np.int64)
    orc.orc_sweep_density(g.sx, g.sy, g.sz, out, g.n_slots, g.b2s, act, act.size, stencil, dt, it, clamp01)
    return out


def test_h_11_density_dequant(pkg, orc):
    g = pkg.SparseGrid("dens", 32, 16, 16, 0.0, 1.0)
    g.ensure_blocks_parallel(np.arange(g.n_blocks))
    for x in range(32):
        g.set_voxel(x, 0, 0, float(np.floor((x / 31.0) * 65535.0 + 0.5)) / 65535.0)
    q = density_of(g)
    f = (q.astype(np.float32) * np.float32(1.0 / 65535.0)).astype(np.float32)
    g2 = g.copy()
    g2._slab[: g2.n_slots] = f
    tile = orc.halo_fill(g2, 0, 1, True)
    assert abs(tile[1, 1, 16] - 15.0 / 31.0) < 1e-3


@pytest.mark.parametrize("stencil", [0, 1, 2])
def test_constant_density_is_a_fixed_point(pkg, orc, stencil):
    g = pkg.SparseGrid("c", 40, 33, 48, 0.0, 1.0)
    g.ensure_blocks_parallel(np.arange(g.n_blocks))
    slab = np.full((g.n_slots, 4096), 40000, np.uint16)
    out = sweep16(orc, g, slab, np.arange(g.n_blocks), stencil, 0.05, it=2)
    assert np.array_equal(out, slab)


@pytest.mark.parametrize("stencil,dt", [(0, 0.1), (1, 0.02), (2, 0.005)])
def test_density_sweep_tracks_the_fp32_sweep(pkg, orc, stencil, dt):
    g = random_grid(pkg, 40, 33, 48, seed=3 + stencil, lo=0.0, hi=1.0, empty=0.0, full=1.0)
    active = np.nonzero(g.b2s >= 2)[0]
    q = density_of(g)
    got = sweep16(orc, g, q, active, stencil, dt).astype(np.float64) / 65535.0
    g32 = g.copy()
    g32._slab[: g32.n_slots] = q.astype(np.float32) * np.float32(1.0 / 65535.0)
    orc.sweep(g32, active, stencil, dt, 1, clamp01=True)
    d = np.abs(got - np.asarray(g32.slab, np.float64))[g.b2s[active]]
    # one store rounding per cell plus neighbours read after their own rounding
    assert d.max() <= 4.0 / 65535.0, d.max() * 65535
