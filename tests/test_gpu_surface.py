"""GPU parity of the surface-reconstruction tail and the block topology,
through the C ABI, against the oracle (bit for bit):

- msbg_surface_place vs orc_surface_place (positions, block ids, order);
- msbg_density_finalize vs orc_density_finalize on every u16 code, and the
  reference's finalize_01..05 known answers on the device;
- reconstruct_and_smooth (place, splat, finalize, mean-curvature sweeps) vs
  the oracle's pipeline;
- msbg_block_topology vs orc_block_topology, and test_ref_02 / test_ref_04.
"""
import numpy as np
import pytest

from test_oracle_surface import EXISTS, finalize as oracle_finalize, finalize_grid, get_voxel, set_voxel, topology

pytestmark = pytest.mark.gpu


def sphere_cloud(n, seed):
    """n points on a unit sphere with 5% radial jitter (LCG-free numpy draw)."""
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    v *= 1.0 + 0.05 * rng.uniform(-1, 1, size=(n, 1))
    return v.astype(np.float32)


def oracle_place(orc, base, cfg):
    from msbg_rs_amd.surface import bounding_box
    bmin, span = bounding_box(base)
    n_inst = cfg.n_instances or base.shape[0]
    cap = n_inst * base.shape[0]
    pos = np.zeros((cap, 3), np.float32)
    bids = np.zeros(cap, np.int64)
    k = orc.lib.orc_surface_place(np.ascontiguousarray(base).reshape(-1), base.shape[0], bmin, span, cfg.sx, cfg.sy,
                                  cfg.sz, cfg.n_instances, cfg.instance_scale_factor, pos.reshape(-1), bids, cap)
    return pos[:k], bids[:k]


@pytest.mark.parametrize("n,s,n_inst,factor", [(2, (32, 32, 32), 0, 0.5), (500, (64, 64, 64), 0, 0.05),
                                               (3000, (96, 64, 80), 64, 0.02), (1500, (40, 33, 48), 300, 0.3)])
def test_place_bitexact(pkg, orc, hip, n, s, n_inst, factor):
    from msbg_rs_amd import surface as S
    base = sphere_cloud(n, n) if n > 2 else np.array([[0, 0, 0], [1, 1, 1]], np.float32)
    cfg = S.SurfaceConfig(*s, n_instances=n_inst, instance_scale_factor=factor)
    bmin, span = S.bounding_box(base)
    placed = S.place(base, bmin, span, cfg, lib=hip)
    want_pos, want_bids = oracle_place(orc, base, cfg)
    got = placed.positions.get()
    assert got.shape == want_pos.shape and got.shape[0] > 0
    assert np.array_equal(got.view(np.uint32), want_pos.view(np.uint32))
    assert np.array_equal(placed.bids.get().astype(np.int64), want_bids)
    assert np.array_equal(placed.active, orc.active_blocks(*s, want_pos, cfg.r_scan()))


def test_place_rejects_too_many_instances(pkg, hip):
    from msbg_rs_amd import surface as S
    from msbg_rs_amd.lib import MsbgError
    base = sphere_cloud(10, 1)
    with pytest.raises(MsbgError):
        S.place(base, *S.bounding_box(base), S.SurfaceConfig(32, 32, 32, n_instances=11), lib=hip)


def device_finalize(pkg, hip, g, slab, active, r_p=2.0, nb=2.0):
    from msbg_rs_amd.particles import DensityGrid
    dg = DensityGrid(g.sx, g.sy, g.sz, capacity_blocks=g.n_slots, lib=hip)
    dg.layout.ensure_blocks_parallel(np.nonzero(g.b2s >= 2)[0])
    b2s_dev = dg.layout.to_host().b2s
    assert np.array_equal(b2s_dev, g.b2s)
    dg.slab.set(np.concatenate([slab, np.zeros((dg.slots - slab.shape[0], 4096), np.uint16)]))
    a = np.ascontiguousarray(active, np.int64)
    hip.call("msbg_density_finalize", dg.layout.h, dg.slab.ptr, dg.slots, a.ctypes.data, a.size, r_p, nb)
    return dg.slab.get()[: slab.shape[0]]


@pytest.mark.parametrize("r_p,nb", [(2.0, 2.0), (1.5, 3.0), (2.5, 1.0)])
def test_finalize_every_code_bitexact(pkg, orc, hip, r_p, nb):
    g, slab = finalize_grid(pkg, (64, 64, 16))
    slab[2:18] = np.arange(65536, dtype=np.uint32).astype(np.uint16).reshape(16, 4096)
    act = np.array([0, 3, 5, 7, 1, 2, 4, 6, 8, 9, 10, 11, 12, 13, 14, 15])
    got = device_finalize(pkg, hip, g, slab, act[:13], r_p, nb)
    want = oracle_finalize(orc, g, slab, act[:13], r_p, nb)
    assert np.array_equal(got, want)  # blocks 13..15 untouched on both sides


def test_finalize_known_answers_on_device(pkg, hip):
    g, slab = finalize_grid(pkg, (32, 32, 32))
    slab[2:] = 65535
    set_voxel(g, slab, 8, 8, 8, 0)
    set_voxel(g, slab, 6, 8, 8, round(2.5 * 2.5 / 16.0 * 65535.0))
    set_voxel(g, slab, 4, 8, 8, 65468)
    set_voxel(g, slab, 20, 20, 20, round(0.9989 * 65535.0))
    set_voxel(g, slab, 21, 20, 20, round(0.9991 * 65535.0))
    out = device_finalize(pkg, hip, g, slab, np.arange(8))
    assert abs(get_voxel(g, out, 8, 8, 8) - 1.0) < 1e-4                                 # finalize_01
    assert abs(get_voxel(g, out, 6, 8, 8) - (1.0 - (0.5 + 2.0) / 6.0)) < 1e-4
    ratio = np.float32(65468.0) / np.float32(65535.0)
    assert abs(get_voxel(g, out, 4, 8, 8) - (1.0 - float(np.sqrt(ratio * 16.0)) / 6.0)) < 1e-4
    assert get_voxel(g, out, 3, 3, 3) == 0.0                                            # finalize_02
    assert get_voxel(g, out, 20, 20, 20) > 0.0 and get_voxel(g, out, 21, 20, 20) == 0.0  # finalize_03
    assert np.array_equal(device_finalize(pkg, hip, g, slab, []), slab)                 # finalize_05


def oracle_reconstruct_and_smooth(pkg, orc, base, cfg):
    from test_oracle_splat import density_grid, run
    import ctypes as C
    fn = orc.lib.orc_density_splat
    fn.restype = None
    fn.argtypes = [C.c_int64] * 3 + [np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS"),
                                     np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS"),
                                     np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS"), C.c_int64, C.c_float,
                                     C.c_int]
    pos, _ = oracle_place(orc, base, cfg)
    s = (cfg.sx, cfg.sy, cfg.sz)
    g, slab = density_grid(pkg, orc, s, pos, cfg.r_scan())
    slab = run(fn, g, slab, pos, cfg.r_scan())
    act = np.nonzero(g.b2s >= 2)[0].astype(np.int64)
    slab = oracle_finalize(orc, g, slab, act, cfg.r_particle, cfg.nb_dist)
    if cfg.n_smooth_iters:
        orc.orc_sweep_density(*s, slab.reshape(-1), g.n_slots, g.b2s, act, act.size, 1, cfg.smooth_dt,
                              cfg.n_smooth_iters, 1)
    return g, slab, act


@pytest.mark.parametrize("res,n,iters", [(64, 400, 0), (64, 400, 5), (96, 900, 3)])
def test_reconstruct_and_smooth_bitexact(pkg, orc, hip, res, n, iters):
    from msbg_rs_amd import surface as S
    base = sphere_cloud(n, res + n)
    cfg = S.SurfaceConfig.demo_case1(res)
    cfg.instance_scale_factor = 0.08
    cfg.n_smooth_iters = iters
    dg, active = S.reconstruct_and_smooth(cfg, base, lib=hip)
    g, want, act = oracle_reconstruct_and_smooth(pkg, orc, base, cfg)
    assert np.array_equal(active, act)
    b2s, got = dg.to_host()
    assert np.array_equal(b2s >= 2, g.b2s >= 2)
    assert np.array_equal(got[b2s[act]], want[g.b2s[act]])
    d = got[b2s[act]].astype(np.float64) / 65535.0
    assert d.max() > 0.5 and d.min() == 0.0  # near the particles and outside every footprint


@pytest.mark.parametrize("dims,n_levels,seed", [((4, 4, 4), 3, 0), ((9, 7, 6), 3, 1), ((20, 17, 23), 4, 2),
                                                ((33, 2, 5), 5, 3), ((1, 1, 1), 2, 4)])
def test_block_topology_bitexact(pkg, orc, hip, dims, n_levels, seed):
    nx, ny, nz = dims
    rng = np.random.default_rng(seed)
    lv = rng.integers(0, n_levels, size=nx * ny * nz).astype(np.uint8)
    orc.orc_regularize(lv, nx, ny, nz, n_levels)
    user = ((rng.random(lv.size) < 0.6).astype(np.uint16) * EXISTS) | (rng.integers(0, 2, lv.size) << 9).astype(
        np.uint16)
    want_flags, want_lists, want_act = topology(orc, lv, dims, n_levels, user)
    for u in (user, None):
        flags, topo = pkg.compute_block_topology(hip.to_device(lv), nx, ny, nz, n_levels,
                                                 None if u is None else hip.to_device(u))
        wf, wl, wa = (want_flags, want_lists, want_act) if u is not None else topology(orc, lv, dims, n_levels)
        assert np.array_equal(flags.get(), wf)
        assert [list(x) for x in topo.blocks_fine_coarse] == [list(x) for x in wl]
        assert topo.n_act_blocks == wa


def test_ref_02_and_04_on_device(pkg, hip):
    lv = np.ones(64, np.uint8)
    lv[21] = 0
    flags, topo = pkg.compute_block_topology(hip.to_device(lv), 4, 4, 4, 3)
    f = flags.get()
    assert f[21] & (1 << 3) and not f[21] & (1 << 2)
    assert list(topo.blocks_fine_coarse[0]) == [21]
    assert not f[63] & ((1 << 3) | (1 << 2))
    flags, _ = pkg.compute_block_topology(hip.to_device(np.zeros(64, np.uint8)), 4, 4, 4, 3)
    f = flags.get()
    assert f[0] & 2 and f[63] & 2 and not f[21] & 2
