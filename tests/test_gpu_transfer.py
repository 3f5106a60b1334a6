"""GPU parity of the transfers: the reference sampler, MAC P2G and the
B-spline P2G / G2P against the oracle."""
import numpy as np
import pytest

import test_oracle_goldens as G
from msbg_rs_amd import transfer as T
from test_gpu_sweep import assert_bitexact, random_grid
from test_oracle_transfer import mac_channels
from transfer_helpers import (dense_from_slabs, oracle_g2p_bspline, oracle_p2g_bspline, oracle_p2g_mac,
                              oracle_p2g_mac_f64)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("interp", [0, 1])
@pytest.mark.parametrize("align", [0, 1])
@pytest.mark.parametrize("bc", [0, 1, 2])
def test_sample_matches_oracle(pkg, orc, hip, interp, align, bc):
    g = random_grid(pkg, 40, 33, 48, seed=40 + interp + 2 * align + 4 * bc, frac=0.8)
    dev = pkg.DeviceGrid.from_host(g, lib=hip)
    rng = np.random.default_rng(bc)
    pos = rng.uniform(-3.0, [43.0, 36.0, 51.0], size=(20000, 3)).astype(np.float32)
    pos[:4] = [[0, 0, 0], [15.5, 15.5, 15.5], [16, 16, 16], [39.9, 32.9, 47.9]]
    got = T.sample(dev, pos, interp, align, bc).get()
    want, _, _ = orc.sample(g, pos, interp, align, bc)
    assert_bitexact(got, want, "sample")


def test_sample_interp_golden_on_device(pkg, hip):
    d = G.GOLD["difftest_interp"]
    g = pkg.SparseGrid("interp", 48, 48, 48, 0.0, 1.0)
    g.fill_from_function(np.arange(g.n_blocks), G.interp_field)
    dev = pkg.DeviceGrid.from_host(g, lib=hip)
    pos = np.array(d["positions"], np.float32)
    gold = np.array(d["golden"], np.float32)
    lin = T.sample(dev, pos, 0, 0, 0).get()
    cub = T.sample(dev, pos, 1, 0, 0).get()
    np.testing.assert_allclose(lin, gold[:, 0], atol=1e-5)
    np.testing.assert_allclose(cub, gold[:, 4], atol=1e-5)


def mac_from_slabs(pkg, hip, b2s, slabs, s):
    mac = T.MacGrid(*s, capacity_blocks=slabs.shape[1], lib=hip)
    for a in range(3):
        host = pkg.SparseGrid("v", *s, 0.0, 0.0)
        host.b2s = b2s.astype(np.int32)
        host._slab = np.ascontiguousarray(slabs[a])
        host.n_slots = slabs.shape[1]
        hip.call("msbg_grid_load", mac.velocity[a].h, host.b2s.ctypes.data, host._slab.ctypes.data, host.n_slots)
    return mac


def test_mac_sampler_and_blend(pkg, orc, hip):
    from test_oracle_transfer import face_field_slabs
    b2s, slabs = face_field_slabs()
    mac = mac_from_slabs(pkg, hip, b2s, slabs, (32, 32, 32))
    got = T.MacSampler(mac, pkg.BoundaryCondition.Clamp).sample(np.array([[5.0, 8.5, 8.5]], np.float32)).get()
    assert list(got[0]) == [5.0, 8.5, 8.5]  # g2p_11
    rng = np.random.default_rng(1)
    pos = rng.uniform(-2, 34, size=(5000, 3)).astype(np.float32)
    for interp in (0, 1):
        got = T.MacSampler(mac, pkg.BoundaryCondition.Clamp).sample(pos, interp).get()
        want = np.zeros((5000, 3), np.float32)
        orc.orc_mac_sample(32, 32, 32, slabs, slabs.shape[1], b2s, pos, 5000, interp, 0, want)
        assert_bitexact(got, want, "mac sample")
    uo, u, ut = (rng.normal(size=(5000, 3)).astype(np.float32) for _ in range(3))
    for alpha in (0.0, 0.95, 1.0):
        got = T.g2p_blend(uo, u, ut, alpha, lib=hip).get()
        want = np.zeros_like(uo)
        orc.orc_g2p_blend(uo, u, ut, 5000, alpha, want)
        assert_bitexact(got, want, "blend")


MAC_TOL = 1e-6  # relative, against the float64-accumulated oracle


def check_mac_against_f64(orc, mac, s, pos, vel, mass, r_p, tol=MAC_TOL):
    """Every face channel of the device splat against the oracle's float64
    sums of the same float32 products: |M - M64| <= tol M64 and |V - V64| <=
    tol (sum |m w u| / M64 + |V64|), V64 = P64 / M64.  Returns the largest
    error over its bound (<= 1)."""
    active, b2s, sums = oracle_p2g_mac_f64(orc, s, pos, vel, mass, r_p)
    absu = oracle_p2g_mac_f64(orc, s, pos, np.abs(vel), mass, r_p)[2]
    worst = 0.0
    for a in range(3):
        m_dev, v_dev = mac.mass[a].to_host(), mac.velocity[a].to_host()
        assert np.array_equal(m_dev.b2s >= 2, b2s >= 2)
        M = m_dev.to_dense().astype(np.float64)
        M64 = dense_from_slabs(b2s, sums[3 + a], s)
        P64 = dense_from_slabs(b2s, sums[a], s)
        Pabs = dense_from_slabs(b2s, absu[a], s)
        em = np.abs(M - M64) / (tol * M64 + 1e-300)
        pos_m = M64 > 0
        V = v_dev.to_dense().astype(np.float64)
        V64 = np.where(pos_m, P64 / np.where(pos_m, M64, 1.0), 0.0)
        ev = np.abs(V - V64) / (tol * np.where(pos_m, Pabs / np.where(pos_m, M64, 1.0) + np.abs(V64), 0.0) + 1e-300)
        worst = max(worst, em.max(), ev.max())
        assert em.max() <= 1.0, ("mass", a, em.max())
        assert ev.max() <= 1.0, ("velocity", a, ev.max())
    return active, worst


@pytest.mark.parametrize("n,s,r_p", [(1, (32, 32, 32), 2.0), (3000, (48, 48, 48), 2.0), (20000, (64, 48, 80), 1.5),
                                     (4000, (40, 33, 20), 2.0), (5000, (48, 48, 48), 3.0)])
def test_particles_to_grid_mac_matches_oracle(pkg, orc, hip, n, s, r_p):
    """The device sums in exact fixed point and rounds a cell at most once
    per committing window, so it is within 1e-6 of the exact sums."""
    rng = np.random.default_rng(n)
    pos = rng.uniform(0.0, np.array(s) - 1e-3, size=(n, 3)).astype(np.float32)  # inside [0, s)
    if n == 1:
        pos[0] = [15.9, 8.0, 8.0]
    vel = rng.normal(size=(n, 3)).astype(np.float32)
    mass = rng.uniform(0.5, 2.0, n).astype(np.float32)
    mac = T.MacGrid(*s, capacity_blocks=((s[0] + 15) // 16) * ((s[1] + 15) // 16) * ((s[2] + 15) // 16) + 8,
                    lib=hip)
    k = T.particles_to_grid(mac, pos, vel, mass, r_p)
    active, worst = check_mac_against_f64(orc, mac, s, pos, vel, mass, r_p)
    assert k == active.size
    print("MAC P2G n=%d: worst error / (1e-6 bound) = %.3f" % (n, worst))


@pytest.mark.parametrize("r_p", [2.0, 3.0])
def test_particles_to_grid_mac_grouped_input(pkg, orc, hip, r_p):
    """Particles in Morton order (each block's particles one run, as a FLIP
    step passes them) take the path without the bucket sort: the same 1e-6
    bound against the float64 sums, and the same result as the shuffled input
    up to the sums' last-bit rounding."""
    s = (48, 40, 56)
    rng = np.random.default_rng(31)
    n = 20000
    pos = rng.uniform(0.0, np.array(s) - 1e-3, size=(n, 3)).astype(np.float32)
    pos[:300] = pos[7]  # a run of equal positions
    vel = rng.normal(size=(n, 3)).astype(np.float32)
    mass = rng.uniform(0.5, 2.0, n).astype(np.float32)
    keys = np.zeros(n, np.uint32)
    orc.orc_morton_keys(np.ascontiguousarray(pos), n, keys)
    order = np.argsort(keys, kind="stable")
    cap = ((s[0] + 15) // 16) * ((s[1] + 15) // 16) * ((s[2] + 15) // 16) + 8
    sorted_mac = T.MacGrid(*s, capacity_blocks=cap, lib=hip)
    k = T.particles_to_grid(sorted_mac, pos[order], vel[order], mass[order], r_p)
    active, worst = check_mac_against_f64(orc, sorted_mac, s, pos[order], vel[order], mass[order], r_p)
    assert k == active.size
    shuffled_mac = T.MacGrid(*s, capacity_blocks=cap, lib=hip)
    T.particles_to_grid(shuffled_mac, pos, vel, mass, r_p)
    for a in range(3):
        ms, mu = sorted_mac.mass[a].to_host().to_dense(), shuffled_mac.mass[a].to_host().to_dense()
        np.testing.assert_allclose(ms, mu, rtol=1e-6, atol=0)
    print("MAC P2G grouped input r_p=%.1f: worst error / (1e-6 bound) = %.3f" % (r_p, worst))


def test_particles_to_grid_mac_dense_cluster(pkg, orc, hip):
    """30,000 particles in a sphere of radius 4 cells (as in a droplet): the
    block holding them is split into work items of 4,096 particles whose
    windows commit with float atomics, up to ~3,700 positive mass terms per
    cell.  Each item's window sum is exact; the cell is rounded once per
    committing item, so it stays within 1e-6 of the exact sum."""
    s = (48, 48, 48)
    rng = np.random.default_rng(7)
    n = 30000
    d = rng.normal(size=(n, 3))
    d *= (4.0 * rng.uniform(0, 1, (n, 1)) ** (1 / 3)) / np.linalg.norm(d, axis=1, keepdims=True)
    pos = (np.array([20.3, 19.7, 21.1]) + d).astype(np.float32)
    vel = rng.normal(size=(n, 3)).astype(np.float32) + np.float32(2.0)
    mass = rng.uniform(0.5, 2.0, n).astype(np.float32)
    mac = T.MacGrid(*s, capacity_blocks=27 + 8, lib=hip)
    k = T.particles_to_grid(mac, pos, vel, mass, 2.0)
    active, worst = check_mac_against_f64(orc, mac, s, pos, vel, mass, 2.0)
    assert k == active.size
    assert mac.mass[0].to_host().to_dense().max() > 500.0  # the dense cells were reached
    print("MAC P2G dense cluster: worst error / (1e-6 bound) = %.3f" % worst)


def test_particles_to_grid_mac_nonfinite_velocity(pkg, orc, hip):
    """One particle with a NaN velocity and one with an infinite one: their
    momentum footprints turn non-finite, every mass cell and every momentum
    cell outside those footprints keeps the exact result of the finite
    particles (the same work items hold all of them)."""
    s = (48, 48, 48)
    rng = np.random.default_rng(3)
    n = 3000
    pos = rng.uniform(0.0, 48 - 1e-3, size=(n, 3)).astype(np.float32)
    pos[:2] = [[20.2, 21.3, 19.8], [30.6, 9.1, 40.2]]
    vel = rng.normal(size=(n, 3)).astype(np.float32)
    vel[0, 0] = np.nan
    vel[1, 2] = np.inf
    mass = rng.uniform(0.5, 2.0, n).astype(np.float32)
    mac = T.MacGrid(*s, capacity_blocks=27 * 4, lib=hip)
    T.particles_to_grid(mac, pos, vel, mass, 2.0)
    fin = vel.copy()
    fin[0, 0] = 0.0
    fin[1, 2] = 0.0
    active, b2s, sums = oracle_p2g_mac_f64(orc, s, pos, fin, mass, 2.0)
    for a in range(3):
        M = mac.mass[a].to_host().to_dense().astype(np.float64)
        M64 = dense_from_slabs(b2s, sums[3 + a], s)
        assert np.all(np.isfinite(M)) and np.all(np.abs(M - M64) <= MAC_TOL * M64 + 1e-300), a
        V = mac.velocity[a].to_host().to_dense()
        bad = ~np.isfinite(V)
        # the non-finite particle's own footprint box of face a
        box = np.zeros_like(bad)
        for p, ch in ((pos[0], 0), (pos[1], 2)):
            if ch != a:
                continue
            of = np.array([1.0 if a == k else 0.5 for k in range(3)])
            lo = np.maximum(np.ceil(p - 2.0 - of).astype(int), 0)
            hi = np.minimum(np.floor(p + 2.0 - of).astype(int), 47)
            box[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = True
        assert not (bad & ~box).any(), a
        if a in (0, 2):
            assert bad.sum() > 20, a
        else:
            assert not bad.any()
        P64 = dense_from_slabs(b2s, sums[a], s)
        ok = ~box & (M64 > 0)
        V64 = P64[ok] / M64[ok]
        assert np.all(np.abs(V[ok] - V64) <= 1e-5 * (np.abs(V64) + 1.0)), a


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("n,s", [(1, (32, 32, 32)), (20000, (64, 48, 80)), (200000, (128, 128, 128))])
def test_bspline_p2g_g2p_bitexact(pkg, orc, hip, order, n, s):
    rng = np.random.default_rng(n + order)
    pos = rng.uniform(0.0, np.array(s) - 1e-3, size=(n, 3)).astype(np.float32)  # inside [0, s)
    if n > 100:
        pos[50:80] = pos[10]  # several particles in one voxel
    vel = rng.normal(size=(n, 3)).astype(np.float32)
    mass = rng.uniform(0.5, 2.0, n).astype(np.float32) if n != 20000 else None
    active, b2s, slabs = oracle_p2g_bspline(orc, order, s, pos, vel, mass)
    tg = T.TransferGrid(*s, capacity_blocks=active.size + 4, lib=hip)
    assert tg.p2g(order, pos, vel, mass) == active.size
    got_m = tg.mass.to_host()
    assert np.array_equal(got_m.b2s >= 2, b2s >= 2)
    assert_bitexact(got_m.to_dense(), dense_from_slabs(b2s, slabs[0], s), "mass")
    for a in range(3):
        assert_bitexact(tg.velocity[a].to_host().to_dense(), dense_from_slabs(b2s, slabs[1 + a], s), "vel %d" % a)
    out = tg.g2p(order, pos).get()
    want = oracle_g2p_bspline(orc, order, s, b2s, slabs[1:], pos)
    assert_bitexact(out, want, "g2p")


@pytest.mark.parametrize("order", [1, 2])
def test_bspline_p2g_dense_plane_bitexact(pkg, orc, hip, order):
    """A voxel plane of a block's neighbourhood holding more particles than the
    P2G stages in LDS (1,024) is read from memory directly; the planes around
    it are staged.  Both give the oracle's bits."""
    s = (48, 48, 48)
    rng = np.random.default_rng(31 + order)
    pos = rng.uniform(0.0, np.array(s) - 1e-3, size=(6000, 3)).astype(np.float32)
    pos[:2500, :2] = rng.uniform(20.0, 23.0, size=(2500, 2))  # a 3 x 3 voxel patch ...
    pos[:2500, 2] = rng.uniform(17.0, 18.0, size=2500)  # ... in one z plane
    vel = rng.normal(size=(6000, 3)).astype(np.float32)
    mass = rng.uniform(0.5, 2.0, 6000).astype(np.float32)
    active, b2s, slabs = oracle_p2g_bspline(orc, order, s, pos, vel, mass)
    tg = T.TransferGrid(*s, capacity_blocks=active.size + 4, lib=hip)
    assert tg.p2g(order, pos, vel, mass) == active.size
    assert_bitexact(tg.mass.to_host().to_dense(), dense_from_slabs(b2s, slabs[0], s), "mass")
    for a in range(3):
        assert_bitexact(tg.velocity[a].to_host().to_dense(), dense_from_slabs(b2s, slabs[1 + a], s), "vel %d" % a)


@pytest.mark.parametrize("order", [1, 2])
def test_bspline_g2p_staged_and_direct_chunks(pkg, orc, hip, order):
    """G2P chunks of 256 sorted particles whose support box fits the LDS
    window (a dense cluster) and chunks whose box does not (sparse particles
    over the whole domain), with particles partly or wholly outside it."""
    s = (96, 80, 112)
    rng = np.random.default_rng(7 + order)
    src = rng.uniform(0.0, np.array(s) - 1e-3, size=(60000, 3)).astype(np.float32)
    vel = rng.normal(size=src.shape).astype(np.float32)
    active, b2s, slabs = oracle_p2g_bspline(orc, order, s, src, vel, None)
    tg = T.TransferGrid(*s, capacity_blocks=active.size + 4, lib=hip)
    assert tg.p2g(order, src, vel, None) == active.size
    cluster = rng.normal(loc=[40.0, 30.0, 50.0], scale=3.0, size=(5000, 3))
    sparse = rng.uniform(-4.0, np.array(s) + 4.0, size=(3000, 3))
    pos = np.concatenate([cluster, sparse]).astype(np.float32)
    rng.shuffle(pos)
    out = tg.g2p(order, pos).get()
    want = oracle_g2p_bspline(orc, order, s, b2s, slabs[1:], pos)
    assert np.any(want[:, 0] == 0.0) and np.any(want[:, 0] != 0.0)
    assert_bitexact(out, want, "g2p")


@pytest.mark.parametrize("r_scan,s,n", [(4.0, (32, 32, 32), 1), (4.0, (48, 40, 33), 3000), (2.5, (40, 33, 48), 5000),
                                        (6.0, (64, 64, 48), 2000), (4.0, (128, 128, 128), 200000)])
def test_density_splat_bitexact(pkg, orc, hip, r_scan, s, n):
    """particles::splat for Density against the oracle, bit for bit (min is
    order-independent): known answer splat_01 for one particle, random clouds
    with particles on block faces, partial domains, margins 3..6."""
    from msbg_rs_amd.particles import DensityGrid
    from test_oracle_splat import density_grid, run
    import ctypes as C
    fn = orc.lib.orc_density_splat
    fn.restype = None
    fn.argtypes = [C.c_int64] * 3 + [np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS"),
                                     np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS"),
                                     np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS"), C.c_int64, C.c_float,
                                     C.c_int]
    rng = np.random.default_rng(n)
    pos = rng.uniform(0.0, np.array(s) - 1e-3, size=(n, 3)).astype(np.float32)
    if n == 1:
        pos[0] = [8.0, 8.0, 8.0]
    else:
        pos[: n // 10] = np.clip(np.round(pos[: n // 10] / 16) * 16 + rng.uniform(-0.6, 0.6, size=(n // 10, 3)),
                                 0.0, np.array(s) - 1e-3)
        pos = pos.astype(np.float32)
    g, slab = density_grid(pkg, orc, s, pos, r_scan)
    want = run(fn, g, slab, pos, r_scan)
    dg = DensityGrid(*s, capacity_blocks=g.n_slots + 4, lib=hip)
    k = dg.splat(pos, r_scan)
    b2s, got = dg.to_host()
    assert k == g.n_slots - 2
    assert np.array_equal(b2s >= 2, g.b2s >= 2)
    act = np.nonzero(g.b2s >= 2)[0]
    assert np.array_equal(got[b2s[act]], want[g.b2s[act]])
    if n == 1:
        assert got[b2s[0], 8 | (8 << 4) | (8 << 8)] == 3072


def test_density_splat_dense_block_bitexact(pkg, orc, hip):
    """A point-cloud instance packed into one block (7,000 particles within 3
    cells of a point, as the surface benchmark places its instances): the
    block's bucket is split into several work items whose windows commit with
    an atomic 16-bit min, and the result is still the oracle's bit for bit."""
    from msbg_rs_amd.particles import DensityGrid
    from test_oracle_splat import density_grid, run
    import ctypes as C
    fn = orc.lib.orc_density_splat
    fn.restype = None
    fn.argtypes = [C.c_int64] * 3 + [np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS"),
                                     np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS"),
                                     np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS"), C.c_int64, C.c_float,
                                     C.c_int]
    s, r_scan = (48, 48, 48), 4.0
    rng = np.random.default_rng(77)
    d = rng.normal(size=(7000, 3))
    d *= (3.0 * rng.uniform(0, 1, (7000, 1)) ** (1 / 3)) / np.linalg.norm(d, axis=1, keepdims=True)
    cluster = np.array([23.7, 24.2, 22.9]) + d
    sparse = rng.uniform(0.0, np.array(s) - 1e-3, size=(1500, 3))
    pos = np.concatenate([cluster, sparse]).astype(np.float32)
    rng.shuffle(pos)
    g, slab = density_grid(pkg, orc, s, pos, r_scan)
    want = run(fn, g, slab, pos, r_scan)
    dg = DensityGrid(*s, capacity_blocks=g.n_slots + 4, lib=hip)
    k = dg.splat(pos, r_scan)
    b2s, got = dg.to_host()
    assert k == g.n_slots - 2
    act = np.nonzero(g.b2s >= 2)[0]
    assert np.array_equal(got[b2s[act]], want[g.b2s[act]])


@pytest.mark.parametrize("s,frac", [((32, 32, 32), 1.0), ((40, 33, 48), 0.7), ((64, 48, 80), 0.5)])
def test_mac_divergence_bitexact(pkg, orc, hip, s, frac):
    from test_oracle_transfer import mac_channels, oracle_divergence
    f = lambda a, X, Y, Z: np.sin(0.37 * X + 0.2 * a) * np.cos(0.11 * Y + a) + 0.05 * Z * (a + 1)
    chans = mac_channels(pkg, s, f, frac, seed=int(frac * 10))
    rng = np.random.default_rng(1)
    act = np.sort(rng.choice(chans[0].n_blocks, size=max(1, chans[0].n_blocks * 2 // 3), replace=False))
    want = oracle_divergence(pkg, orc, chans, act)
    mac = T.MacGrid(*s, capacity_blocks=chans[0].n_blocks + 4, lib=hip)
    slabs = [np.ascontiguousarray(c.slab) for c in chans]  # alive for the copies
    for a in range(3):
        hip.call("msbg_grid_load", mac.velocity[a].h, chans[a].b2s.ctypes.data, slabs[a].ctypes.data,
                 chans[a].n_slots)
    dst = pkg.DeviceGrid(*s, 4, 0.0, 0.0, capacity_blocks=chans[0].n_blocks + 4, lib=hip)
    mac.divergence(dst, act)
    got = dst.to_host()
    assert np.array_equal(got.b2s >= 2, want.b2s >= 2)
    for b in act:
        assert_bitexact(got.slab[got.b2s[b]], want.slab[want.b2s[b]], "div block %d" % b)


def morton_order(pos):
    """Permutation sorting positions by the Morton code of their cell (cells
    clipped to [0, 1023]): consecutive particles are spatial neighbours."""
    c = np.clip(np.floor(pos), 0, 1023).astype(np.uint64)
    key = np.zeros(pos.shape[0], np.uint64)
    for bit in range(10):
        for d in range(3):
            key |= ((c[:, d] >> np.uint64(bit)) & np.uint64(1)) << np.uint64(3 * bit + d)
    return np.argsort(key, kind="stable")


@pytest.mark.parametrize("interp,bc,n,sort", [(0, 0, 20000, False), (1, 0, 20000, False), (0, 1, 20000, False),
                                              (0, 0, 300000, True), (0, 2, 300000, True), (0, 1, 50000, True)])
def test_g2p_flip_matches_sample_twice_and_blend(pkg, orc, hip, interp, bc, n, sort):
    """msbg_g2p_flip: bit for bit the oracle's two MAC samples and blend, on
    random and on Morton-sorted particles (the order a FLIP step passes),
    some of them near or beyond the domain's faces."""
    s = (40, 33, 48)
    f = lambda a, X, Y, Z: np.sin(0.21 * X + a) * np.cos(0.13 * Y) + 0.03 * Z
    g = lambda a, X, Y, Z: np.cos(0.17 * X) * np.sin(0.09 * Y + a) - 0.02 * Z
    from test_gpu_flip import mac_from_channels
    cu, cut = mac_channels(pkg, s, f), mac_channels(pkg, s, g)  # every block: one slot table for all channels
    mu, mut = mac_from_channels(pkg, hip, cu), mac_from_channels(pkg, hip, cut)
    rng = np.random.default_rng(9 + n)
    pos = rng.uniform(-1.0, np.array(s) + 1.0, size=(n, 3)).astype(np.float32)  # some outside the domain
    if sort:
        pos = pos[morton_order(pos)]
    uo = rng.normal(size=(n, 3)).astype(np.float32)
    for alpha in (0.0, 0.95, 1.0):
        got = T.g2p_flip(mu, mut, pos, uo, alpha, interp, bc).get()
        su, st = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32)
        for c, out in ((cu, su), (cut, st)):
            slabs = np.ascontiguousarray(np.stack([np.ascontiguousarray(x.slab) for x in c]))
            orc.orc_mac_sample(*s, slabs, slabs.shape[1], c[0].b2s, pos, n, interp, bc, out)
        want = np.zeros_like(uo)
        orc.orc_g2p_blend(uo, su, st, n, alpha, want)
        assert_bitexact(got, want, "g2p flip")
