"""GPU parity of the halo gather and the 8-colour sweeps against the oracle.

Every comparison is bit-exact (the kernels repeat the reference's float
operation order), plus the reference's golden samples where they exist.
"""
import numpy as np
import pytest

import test_oracle_goldens as G

pytestmark = pytest.mark.gpu


def random_grid(pkg, sx, sy, sz, seed, frac=0.7, n_full=2, n_empty=1, lo=-1.0, hi=2.0, empty=0.25, full=0.75):
    rng = np.random.default_rng(seed)
    g = pkg.SparseGrid("rand", sx, sy, sz, empty, full)
    bids = np.nonzero(rng.random(g.n_blocks) < frac)[0]
    g.ensure_blocks_parallel(bids)
    g._slab[2:g.n_slots] = rng.uniform(lo, hi, size=(g.n_slots - 2, g.N)).astype(np.float32)
    free = np.nonzero(g.b2s < 2)[0]
    rng.shuffle(free)
    for b in free[:n_full]:
        g.set_full_block(int(b))
    return g


def assert_bitexact(a, b, what):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    same = (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))
    if not same.all():
        i = np.argmax(~same.ravel())
        raise AssertionError("%s: %d of %d values differ; first at %d: %r vs %r (max abs diff %g)" % (
            what, (~same).sum(), same.size, i, a.ravel()[i], b.ravel()[i], np.nanmax(np.abs(a - b))))


def test_cpp_hash_device_roundtrip(pkg, hip):
    g = G.cpp_script_grid(pkg)
    d = pkg.DeviceGrid.from_host(g, lib=hip)
    back = d.to_host()
    assert G.grid_hash(back.to_dense()) == int(G.GOLD["difftest_cpp"]["hash"], 16)


@pytest.mark.parametrize("halo,full", [(1, False), (1, True), (2, True), (2, False)])
@pytest.mark.parametrize("bc", [0, 1, 2])
@pytest.mark.parametrize("shape", [(48, 48, 48), (40, 33, 17), (17, 16, 16), (16, 16, 16)])
def test_halo_gather_matches_oracle(pkg, orc, hip, halo, full, bc, shape):
    g = random_grid(pkg, *shape, seed=hash((halo, full, bc) + shape) & 0xFFFF)
    d = pkg.DeviceGrid.from_host(g, lib=hip)
    bids = np.arange(g.n_blocks)
    tiles = pkg.halo_fill(d, bids, halo, full, bc)
    for i, b in enumerate(bids):
        want = orc.halo_fill(g, int(b), halo, full, bc)
        assert_bitexact(tiles[i], want, "halo bid %d" % b)


def test_halo_known_answers_on_device(pkg, hip):
    g = G.bid_grid(pkg)
    d = pkg.DeviceGrid.from_host(g, lib=hip)
    t = pkg.halo_fill(d, [13], 1, True)[0]
    assert t[1, 1, 1] == 13 and t[1, 1, 0] == 12 and t[1, 1, 17] == 14 and t[17, 17, 17] == 26
    t = pkg.halo_fill(d, [13], 2, True)[0]
    assert t[2, 2, 0] == 12 and t[2, 2, 19] == 14 and t[0, 0, 0] == 0 and t[19, 19, 19] == 26


@pytest.mark.parametrize("name,stencil", [("laplacian", 0), ("meancurv", 1), ("bilaplacian", 2)])
def test_stencil_golden_single_block(pkg, orc, hip, name, stencil):
    """difftest_stencil: one application on block (1,1,1) of the polynomial field."""
    d = G.GOLD["difftest_stencil"]
    g = G.build_stencil_grid(pkg)
    bid = g.get_block_id(16, 16, 16)
    dev = pkg.DeviceGrid.from_host(g, lib=hip)
    pkg.Sweeper(dev).sweep([bid], pkg.Stencil(stencil), pkg.PdeParams(d["dt"], 1, False))
    got = dev.read_blocks([bid])[0]
    np.testing.assert_allclose(got[d["sample_idx"]], d["golden"][name], atol=1e-5, rtol=0)
    halo, full = (2, True) if stencil == 2 else (1, stencil == 1)
    want = orc.kernel(stencil, orc.halo_fill(g, bid, halo, full), d["dt"])
    assert_bitexact(got, want, name)


@pytest.mark.parametrize("name,stencil", [("laplacian", 0), ("meancurv", 1)])
def test_smoother_golden_4_iterations(pkg, orc, hip, name, stencil):
    """difftest_smoother: Sweeper, 4 iterations over the 64^3 sin/cos field."""
    d = G.GOLD["difftest_smoother"]
    g = pkg.SparseGrid("smoothertest", 64, 64, 64, 0.0, 1.0)
    g.fill_from_function(np.arange(g.n_blocks), G.smoother_field)
    dev = pkg.DeviceGrid.from_host(g, lib=hip)
    pkg.Sweeper(dev).sweep(np.arange(64), pkg.Stencil(stencil), pkg.PdeParams(d["dt"], d["iterations"]))
    got = dev.read_blocks(np.arange(64)).ravel()
    np.testing.assert_allclose(got[d["sample_idx"]], d["golden"][name], atol=1e-6, rtol=0)
    orc.sweep(g, np.arange(64), stencil, d["dt"], d["iterations"])
    assert_bitexact(got, np.concatenate([g.block_data(b) for b in range(64)]), name)


@pytest.mark.parametrize("stencil", [0, 1, 2])
@pytest.mark.parametrize("clamp01", [False, True])
@pytest.mark.parametrize("shape,seed", [((48, 48, 48), 1), ((40, 33, 17), 2), ((64, 32, 48), 3)])
def test_sweep_matches_oracle(pkg, orc, hip, stencil, clamp01, shape, seed):
    g = random_grid(pkg, *shape, seed=seed, lo=-0.2, hi=1.2)
    rng = np.random.default_rng(seed)
    # active list: value blocks, a full dummy, an empty block and a duplicate-free shuffle
    active = list(np.nonzero(g.b2s >= 2)[0])
    active += list(np.nonzero(g.b2s == 1)[0][:1]) + list(np.nonzero(g.b2s == 0)[0][:1])
    active = np.array(active, np.int64)
    rng.shuffle(active)
    dev = pkg.DeviceGrid.from_host(g, lib=hip)
    dt = [0.05, 0.02, 0.01][stencil]
    pkg.Sweeper(dev).sweep(active, pkg.Stencil(stencil), pkg.PdeParams(dt, 3, clamp01))
    orc.sweep(g, active, stencil, dt, 3, clamp01)
    back = dev.to_host()
    assert_bitexact(back.slab, g.slab, "stencil %d" % stencil)


def test_sweep_affine_fixed_point(pkg, hip):
    """solver.rs:334-361: an affine field is a fixed point of one Laplacian sweep in the interior."""
    g = pkg.SparseGrid("aff", 32, 32, 32)
    g.fill_from_function(np.arange(g.n_blocks), lambda X, Y, Z: (2.0 * X + 3.0 * Y - Z).astype(np.float32))
    dev = pkg.DeviceGrid.from_host(g, lib=hip)
    pkg.Sweeper(dev).sweep(np.arange(g.n_blocks), pkg.Stencil.Laplacian, pkg.PdeParams(0.05, 1, False))
    out = dev.to_host().to_dense()
    z, y, x = np.meshgrid(np.arange(32), np.arange(32), np.arange(32), indexing="ij")
    want = 2.0 * x + 3.0 * y - z
    assert np.abs(out - want)[1:31, 1:31, 1:31].max() < 1e-4


def test_sweep_empty_active_and_errors(pkg, hip):
    g = pkg.SparseGrid("e", 32, 32, 32)
    dev = pkg.DeviceGrid.from_host(g, lib=hip)
    pkg.Sweeper(dev).sweep([], pkg.Stencil.Laplacian, pkg.PdeParams(0.05, 2, True))
    with pytest.raises(pkg.MsbgError):
        pkg.Sweeper(dev).sweep([10 ** 6], pkg.Stencil.Laplacian, pkg.PdeParams(0.05, 1))
    with pytest.raises(pkg.MsbgError):
        pkg.Sweeper(dev).sweep([0], 7, pkg.PdeParams(0.05, 1))


def test_color_buckets_match_oracle(pkg, orc, hip):
    g = random_grid(pkg, 128, 96, 64, seed=5, frac=0.5)
    active = np.nonzero(g.b2s >= 2)[0]
    dev = pkg.DeviceGrid.from_host(g, lib=hip)
    got, gc = pkg.color_buckets(dev, active)
    want, wc = orc.color_buckets(g, active)
    assert np.array_equal(got, want) and np.array_equal(gc, wc)


def multires_pair(pkg, seed, sx=64, sy=48, sz=32):
    rng = np.random.default_rng(seed)
    fine = pkg.SparseGrid("f", sx, sy, sz, 0.0, 0.0)
    coarse = pkg.SparseGrid("c", sx // 2, sy // 2, sz // 2, 0.0, 0.0, log2_bsx=3)
    levels = rng.integers(0, 2, size=fine.n_blocks).astype(np.uint8)
    fb = np.nonzero(levels == 0)[0]
    cb = np.nonzero(levels == 1)[0]
    fine.ensure_blocks_parallel(fb)
    fine._slab[2:fine.n_slots] = rng.uniform(-1, 1, size=(fine.n_slots - 2, 4096)).astype(np.float32)
    coarse.ensure_blocks_parallel(cb)
    coarse._slab[2:coarse.n_slots] = rng.uniform(-1, 1, size=(coarse.n_slots - 2, 512)).astype(np.float32)
    return fine, coarse, levels


@pytest.mark.parametrize("halo,full", [(1, True), (1, False), (2, True)])
def test_multires_gather_matches_oracle(pkg, orc, hip, halo, full):
    fine, coarse, levels = multires_pair(pkg, 7)
    df, dc = pkg.DeviceGrid.from_host(fine, lib=hip), pkg.DeviceGrid.from_host(coarse, lib=hip)
    dl = hip.to_device(levels)
    bids = np.nonzero(levels == 0)[0]
    tiles = pkg.halo_fill_multires(df, dc, dl, bids, halo, full)
    hsx = 16 + 2 * halo
    for i, b in enumerate(bids):
        want = np.zeros(hsx ** 3, np.float32)
        orc.orc_halo_fill_multires(fine.sx, fine.sy, fine.sz, 4, fine.slab, fine.b2s, coarse.slab, coarse.b2s,
                                   levels, int(b), halo, int(full), 1, want)
        assert_bitexact(tiles[i].ravel(), want, "multires bid %d" % b)


@pytest.mark.parametrize("stencil", [0, 1, 2])
def test_multires_sweep_matches_oracle(pkg, orc, hip, stencil):
    fine, coarse, levels = multires_pair(pkg, 11 + stencil)
    df, dc = pkg.DeviceGrid.from_host(fine, lib=hip), pkg.DeviceGrid.from_host(coarse, lib=hip)
    dl = hip.to_device(levels)
    active = np.nonzero(levels == 0)[0]
    pkg.sweep_multires(df, dc, dl, active, stencil, pkg.PdeParams(0.02, 2, False))
    slab = np.ascontiguousarray(fine.slab)
    orc.orc_sweep_multires(fine.sx, fine.sy, fine.sz, 4, slab, fine.b2s, coarse.slab, coarse.b2s, levels, active,
                           active.size, stencil, 0.02, 2, 0, 1)
    assert_bitexact(df.to_host().slab, slab, "multires sweep")


def level_pair(pkg, orc, hip, s, n_levels, seed, **kw):
    from test_oracle_multires import LevelGrid
    lg = LevelGrid(pkg, orc, s, n_levels, seed=seed, lo=0.0, hi=1.0, **kw)
    dev = [pkg.DeviceGrid.from_host(g, lib=hip) for g in lg.grids]
    return lg, dev, hip.to_device(lg.levels)


@pytest.mark.parametrize("stencil,dt", [(0, 0.1), (1, 0.02), (2, 0.005)])
@pytest.mark.parametrize("s,n_levels", [((160, 128, 128), 3), ((160, 128, 128), 4), ((40, 56, 72), 4)])
def test_level_sweep_matches_oracle(pkg, orc, hip, stencil, dt, s, n_levels):
    """All-level sweeps (every block at its own level; coarse halos read finer
    neighbours as 2^3 means and coarser ones by injection) bit for bit against
    the oracle, including partial blocks at every level ((40, 56, 72))."""
    lg, dev, dl = level_pair(pkg, orc, hip, s, n_levels, seed=stencil + 3 * n_levels + s[0])
    active = np.arange(lg.levels.size)
    pkg.sweep_levels(dev, dl, active, stencil, pkg.PdeParams(dt, 2, stencil == 0))
    lg.sweep(orc, active, stencil, dt, 2, stencil == 0)
    for l in range(n_levels):
        assert_bitexact(dev[l].to_host().slab, lg.slabs[l], "level %d" % l)


def test_level_sweep_graph_alternating(pkg, orc, hip):
    """A plan over a subset of the blocks of all levels, MC19 and BiLap25
    graphs replayed alternately without synchronisation, against the oracle."""
    lg, dev, dl = level_pair(pkg, orc, hip, (160, 128, 128), 4, seed=5)
    rng = np.random.default_rng(2)
    active = np.nonzero(rng.random(lg.levels.size) < 0.8)[0]
    plan = pkg.SweepPlan(dev[0], active, levels=dl, level_grids=dev)
    keys = [(1, 0.02), (2, 0.005)] * 3
    for st, dt in keys:
        plan.run_graph_async(st, dt, 1)
    hip.synchronize()
    for st, dt in keys:
        lg.sweep(orc, active, st, dt, 1)
    for l in range(4):
        assert_bitexact(dev[l].to_host().slab, lg.slabs[l], "level %d" % l)


def test_level_sweep_five_levels(pkg, orc, hip):
    """Five levels: edge-1 blocks at level 4 take the radius-1 stencils; the
    radius-2 bi-Laplacian is refused with an error, and nothing is written."""
    # few fine blocks and few mixed ones, so that 2:1 regularisation leaves a
    # staircase down to level 4 (41 edge-1 blocks here)
    lg, dev, dl = level_pair(pkg, orc, hip, (160, 128, 128), 5, seed=9, frac_l0=0.01, frac_mixed=0.05)
    assert (lg.levels == 4).sum() > 10 and all((lg.levels == l).any() for l in range(5))
    active = np.arange(lg.levels.size)
    for st, dt in ((0, 0.1), (1, 0.02)):
        pkg.sweep_levels(dev, dl, active, st, pkg.PdeParams(dt, 1, False))
        lg.sweep(orc, active, st, dt, 1)
    for l in range(5):
        assert_bitexact(dev[l].to_host().slab, lg.slabs[l], "level %d" % l)
    with pytest.raises(pkg.MsbgError):
        pkg.sweep_levels(dev, dl, active, 2, pkg.PdeParams(0.005, 1, False))
    for l in range(5):
        assert_bitexact(dev[l].to_host().slab, lg.slabs[l], "level %d after the refused call" % l)


def test_downsample_matches_oracle(pkg, orc, hip):
    fine, coarse, levels = multires_pair(pkg, 3)
    bids = np.nonzero(levels == 0)[0]
    df = pkg.DeviceGrid.from_host(fine, lib=hip)
    dc = pkg.DeviceGrid.from_host(coarse, capacity_blocks=coarse.n_blocks, lib=hip)
    pkg.downsample_channel_avg(dc, df, bids)
    got = dc.to_host()
    coarse.ensure_blocks_parallel(bids)
    cs = np.ascontiguousarray(coarse.slab)
    orc.orc_downsample_avg(4, fine.slab, fine.b2s, cs, coarse.b2s, bids, bids.size)
    for b in bids:
        assert_bitexact(got.slab[got.b2s[b]], cs[coarse.b2s[b]], "downsample %d" % b)


def test_regularize_matches_oracle(pkg, orc, hip):
    rng = np.random.default_rng(9)
    nx, ny, nz = 20, 17, 13
    levels = rng.integers(0, 4, size=nx * ny * nz).astype(np.uint8)
    levels[rng.random(levels.size) < 0.8] = 3
    dl = hip.to_device(levels)
    pkg.regularize_refinement_map(dl, nx, ny, nz, 4)
    want = levels.copy()
    orc.orc_regularize(want, nx, ny, nz, 4)
    assert np.array_equal(dl.get(), want)


@pytest.mark.parametrize("halo", [1, 2])
def test_halo_topology_matches_oracle(pkg, orc, hip, halo):
    g = random_grid(pkg, 80, 48, 33, seed=21 + halo, frac=0.5, n_full=3)
    dev = pkg.DeviceGrid.from_host(g, lib=hip)
    bids = np.arange(g.n_blocks)
    got = pkg.halo_topology(dev, bids, halo)
    nb = (2 * halo + 1) ** 3
    for i, b in enumerate(bids):
        want = np.zeros(nb, np.int32)
        orc.orc_halo_topology(g.sx, g.sy, g.sz, 4, g.b2s, None, None, int(b), halo, want)
        assert np.array_equal(got[i], want), b


def test_halo_topology_multires_matches_oracle(pkg, orc, hip):
    fine, coarse, levels = multires_pair(pkg, 17)
    df, dc = pkg.DeviceGrid.from_host(fine, lib=hip), pkg.DeviceGrid.from_host(coarse, lib=hip)
    dl = hip.to_device(levels)
    bids = np.arange(fine.n_blocks)
    for halo in (1, 2):
        got = pkg.halo_topology(df, bids, halo, coarse=dc, levels=dl)
        nb = (2 * halo + 1) ** 3
        for i, b in enumerate(bids):
            want = np.zeros(nb, np.int32)
            orc.orc_halo_topology(fine.sx, fine.sy, fine.sz, 4, fine.b2s, coarse.b2s.ctypes.data,
                                  levels.ctypes.data, int(b), halo, want)
            assert np.array_equal(got[i], want), (halo, b)


@pytest.mark.parametrize("stencil", [0, 1, 2])
def test_graph_replay_matches_eager(pkg, orc, hip, stencil):
    """Each graph replay, with eager work and synchronizations between replays,
    equals the eager launches of the same sweep (bit-exact)."""
    g = random_grid(pkg, 64, 48, 48, seed=31 + stencil, lo=0.0, hi=1.0)
    active = np.nonzero(g.b2s >= 2)[0]
    ga = pkg.DeviceGrid.from_host(g, lib=hip)
    gb = pkg.DeviceGrid.from_host(g, lib=hip)
    pa, pb = pkg.SweepPlan(ga, active), pkg.SweepPlan(gb, active)
    for rep in range(3):
        pa.run_async(stencil, 0.02, 2, True)
        hip.synchronize()
        _ = pkg.halo_fill(gb, active[:3], 1, True)  # eager work between replays
        pb.run_graph_async(stencil, 0.02, 2, True)
        hip.synchronize()
        assert_bitexact(gb.to_host().slab, ga.to_host().slab, "replay %d" % rep)
    orc.sweep(g, active, stencil, 0.02, 6, True)
    assert_bitexact(gb.to_host().slab, g.slab, "graph vs oracle")


@pytest.mark.parametrize("stencil,dt", [(0, 0.1), (1, 0.02), (2, 0.005)])
@pytest.mark.parametrize("s", [(40, 33, 48), (64, 48, 80)])
def test_density8_sweep_bitexact(pkg, orc, hip, stencil, dt, s):
    """Sweeper<Density8>: u8 storage, linear dequantised halo, sqrt-compressed
    stochastic-rounding store with the stated (seed, sweep, bid, voxel)
    stream; a one-shot call (sweeps 0, 1) and a plan run eagerly once then
    replayed as a graph twice (sweeps 0, 1, 2) against the oracle."""
    import ctypes as C
    g = random_grid(pkg, *s, seed=21 + stencil, lo=0.0, hi=1.0, empty=0.0, full=1.0)
    active = np.nonzero(g.b2s >= 2)[0].astype(np.int64)
    rng = np.random.default_rng(stencil)
    q = rng.integers(0, 256, size=(g.n_slots, 4096)).astype(np.uint8)
    q[0], q[1] = 0, 255
    layout = pkg.DeviceGrid.from_host(g, lib=hip)
    want = q.copy()
    orc.orc_sweep_density8(g.sx, g.sy, g.sz, want, g.n_slots, g.b2s, active, active.size, stencil, dt, 2, 1, 77, 0)
    d = hip.to_device(q)
    hip.call("msbg_sweep_density8", layout.h, d.ptr, active.ctypes.data, active.size, stencil, dt, 2, 1, 77)
    assert np.array_equal(d.get(), want)
    want3 = q.copy()
    orc.orc_sweep_density8(g.sx, g.sy, g.sz, want3, g.n_slots, g.b2s, active, active.size, stencil, dt, 3, 1, 78, 0)
    d2 = hip.to_device(q)
    plan = C.c_void_p()
    hip.call("msbg_sweep_plan_new_density8", layout.h, d2.ptr, active.ctypes.data, active.size, 78, C.byref(plan))
    hip.call("msbg_sweep_plan_run_async", plan, stencil, dt, 1, 1)
    hip.call("msbg_sweep_plan_run_graph_async", plan, stencil, dt, 1, 1)
    hip.call("msbg_sweep_plan_run_graph_async", plan, stencil, dt, 1, 1)
    hip.synchronize()
    hip.call("msbg_sweep_plan_free", plan)
    assert np.array_equal(d2.get(), want3)


def test_graph_cache_alternating_stencils_without_sync(pkg, orc, hip):
    """A multires plan replays MC19 and BiLap25 graphs back to back with no
    synchronisation in between (the benchmark's step), then a fifth key forces
    an eviction while earlier graphs may still run: every result equals the
    oracle's sweeps in the same order, bit for bit."""
    fine, coarse, levels = multires_pair(pkg, 23)
    fine._slab[2:] = np.abs(fine._slab[2:])
    df, dc = pkg.DeviceGrid.from_host(fine, lib=hip), pkg.DeviceGrid.from_host(coarse, lib=hip)
    dl = hip.to_device(levels)
    active = np.nonzero(levels == 0)[0]
    plan = pkg.SweepPlan(df, active, coarse=dc, levels=dl)
    keys = [(1, 0.02, 1, False), (2, 0.005, 1, False)] * 4 + [(0, 0.1, 1, True), (1, 0.03, 2, False),
                                                              (2, 0.004, 1, True), (1, 0.02, 1, False)]
    for st, dt, it, cl in keys:
        plan.run_graph_async(st, dt, it, cl)
    hip.synchronize()
    slab = np.ascontiguousarray(fine.slab)
    for st, dt, it, cl in keys:
        orc.orc_sweep_multires(fine.sx, fine.sy, fine.sz, 4, slab, fine.b2s, coarse.slab, coarse.b2s, levels,
                               active, active.size, st, dt, it, int(cl), 1)
    assert_bitexact(df.to_host().slab, slab, "alternating graph replays")


@pytest.mark.parametrize("stencil,dt", [(0, 0.1), (1, 0.02), (2, 0.005)])
@pytest.mark.parametrize("s", [(40, 33, 48), (64, 48, 80)])
def test_density_sweep_bitexact(pkg, orc, hip, stencil, dt, s):
    """Sweeper<Density>: u16 storage, dequantised halo, round-half-away store;
    eager and graph-replayed plans against the oracle, two iterations."""
    import ctypes as C
    from test_oracle_density import density_of, sweep16
    g = random_grid(pkg, *s, seed=11 + stencil, lo=0.0, hi=1.0, empty=0.0, full=1.0)
    active = np.nonzero(g.b2s >= 2)[0].astype(np.int64)
    q = density_of(g)
    q[0], q[1] = 0, 65535
    want = sweep16(orc, g, q, active, stencil, dt, it=2)
    layout = pkg.DeviceGrid.from_host(g, lib=hip)
    d = hip.to_device(q)
    hip.call("msbg_sweep_density", layout.h, d.ptr, active.ctypes.data, active.size, stencil, dt, 2, 1)
    assert np.array_equal(d.get(), want)
    d2 = hip.to_device(q)
    plan = C.c_void_p()
    hip.call("msbg_sweep_plan_new_density", layout.h, d2.ptr, active.ctypes.data, active.size, C.byref(plan))
    hip.call("msbg_sweep_plan_run_graph_async", plan, stencil, dt, 2, 1)
    hip.synchronize()
    hip.call("msbg_sweep_plan_free", plan)
    assert np.array_equal(d2.get(), want)
