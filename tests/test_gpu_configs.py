"""GPU parity at the benchmarked workloads themselves, through the C ABI.

- 128^3 shell (280 blocks): 20 weighted-Jacobi sweeps replayed as one graph
  (FP32 / FP16 / BF16) against 20 oracle Jacobi applications, and the
  trilinear P2G / G2P of the 1M shell particles, bit for bit.
- 512^3 4-level noisy sphere: the benchmark's multires sweep plan (MC19 and
  BiLap25 graphs replayed alternately, no synchronisation in between) against
  the oracle's multires sweeps; the same grid swept at all four levels; the
  16M band particles' Morton sort permutation; quadratic B-spline P2G on a
  sample of blocks (with every particle that reaches them) and G2P on a sample
  of particles, all bit for bit.
- 1024^3 dam break with 64M particles: one FLIP step (FP16 band storage,
  multigrid); the P2G on a sample of blocks within 1e-6 of the float64 sums,
  the staggered G2P + blend on a sample of particles bit for bit, and the
  post-projection divergence equal to the solve's residual.
"""
import ctypes as C

import numpy as np
import pytest

from solve_helpers import oracle_jacobi
from test_gpu_sweep import assert_bitexact
from transfer_helpers import oracle_g2p_bspline, oracle_p2g_bspline, oracle_p2g_mac_f64

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def near_blocks(pos, pick, nb):
    """Particles whose 16^3 block is one of the 27 around a picked block (of
    an nb^3 block grid): a superset of every particle whose transfer footprint
    (at most 4 cells) reaches a picked block."""
    mark = np.zeros(nb ** 3, bool)
    bx, by, bz = pick % nb, (pick // nb) % nb, pick // (nb * nb)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                x, y, z = bx + dx, by + dy, bz + dz
                ok = (x >= 0) & (x < nb) & (y >= 0) & (y < nb) & (z >= 0) & (z < nb)
                mark[(x + nb * (y + nb * z))[ok]] = True
    b = np.minimum((pos // 16).astype(np.int64), nb - 1)
    return mark[b[:, 0] + nb * (b[:, 1] + nb * b[:, 2])]


# ---------------------------------------------------------------------------
# 128^3 shell
@pytest.fixture(scope="module")
def shell():
    from msbg_rs_amd import scenes
    return scenes.shell_scene()


@pytest.mark.parametrize("st", [0, 1, 2])
def test_shell_jacobi20_graph_bitexact(pkg, orc, hip, shell, st):
    """The bench's shell leg: x -> tmp -> x ping-pong of 20 sweeps with b = 0,
    twice in a row without a sync, then on a second pair of buffers (the graph
    is captured again)."""
    from solve_helpers import to_storage
    g, bids, _ = shell
    say("shell jacobi20 storage", st, "blocks", bids.size)
    layout = pkg.DeviceGrid.from_host(g, lib=hip)
    band = pkg.PoissonBand(layout, bids, lib=hip)
    x0 = to_storage(np.ascontiguousarray(g.slab, np.float32), st)
    zero = to_storage(np.zeros_like(np.asarray(g.slab, np.float32)), st)
    want = x0
    for _ in range(40):
        want = oracle_jacobi(orc, g, bids, want, zero, 2.0 / 3.0, st)
    dz = hip.to_device(zero)
    x, tmp = hip.to_device(x0), hip.to_device(x0)
    band.jacobi_sweeps(st, x, tmp, dz, 2.0 / 3.0, 20)
    band.jacobi_sweeps(st, x, tmp, dz, 2.0 / 3.0, 20)
    x2, tmp2 = hip.to_device(x0), hip.to_device(x0)
    band.jacobi_sweeps(st, x2, tmp2, dz, 2.0 / 3.0, 20)
    band.jacobi_sweeps(st, x2, tmp2, dz, 2.0 / 3.0, 20)
    for got in (x.get(), x2.get()):
        assert np.array_equal(np.ascontiguousarray(got).view(np.uint8), np.ascontiguousarray(want).view(np.uint8))


def test_shell_trilinear_p2g_g2p_bitexact(pkg, orc, hip, shell):
    """The 1M shell particles (seeded as the bench's), trilinear P2G and G2P
    over the whole 128^3 domain."""
    from msbg_rs_amd import scenes
    from msbg_rs_amd import transfer as T
    _, bids, centre = shell
    pos, vel = scenes.shell_particles(1 << 20, centre)
    s = (128, 128, 128)
    say("shell trilinear transfers", pos.shape[0])
    active, b2s, slabs = oracle_p2g_bspline(orc, 1, s, pos, vel)
    tg = T.TransferGrid(*s, capacity_blocks=active.size + 4, lib=hip)
    assert tg.p2g(1, pos, vel) == active.size
    m = tg.mass.to_host()
    assert np.array_equal(m.b2s >= 2, b2s >= 2)
    assert_bitexact(m.slab[m.b2s[active]], slabs[0][b2s[active]], "mass")
    for a in range(3):
        v = tg.velocity[a].to_host()
        assert_bitexact(v.slab[v.b2s[active]], slabs[1 + a][b2s[active]], "velocity %d" % a)
    out = tg.g2p(1, pos).get()
    assert_bitexact(out, oracle_g2p_bspline(orc, 1, s, b2s, slabs[1:], pos), "g2p")


# ---------------------------------------------------------------------------
# 512^3 4-level noisy sphere
@pytest.fixture(scope="module")
def sphere(pkg):
    from msbg_rs_amd import scenes
    return scenes.noisy_sphere_scene(S=512, n_levels=4)


def test_sphere512_multires_graph_steps_bitexact(pkg, orc, hip, sphere):
    """The benchmark's step, exactly: a multires plan over the 4288 level-0
    blocks, MC19 (dt 0.02) and BiLap25 (dt 0.005) graphs replayed alternately
    for three steps with no synchronisation in between, against three oracle
    steps of the same multires sweeps."""
    sc = sphere
    fine_h, coarse_h = sc.grids[0], sc.grids[1]
    active = sc.level_blocks(0)
    say("512^3 multires steps over", active.size, "level-0 blocks")
    fine = pkg.DeviceGrid.from_host(fine_h, lib=hip)
    coarse = pkg.DeviceGrid.from_host(coarse_h, lib=hip)
    levels = hip.to_device(sc.levels)
    plan = pkg.SweepPlan(fine, active, coarse=coarse, levels=levels)
    steps = [(1, 0.02), (2, 0.005)] * 3
    for st, dt in steps:
        plan.run_graph_async(st, dt, 1)
    hip.synchronize()
    slab = np.ascontiguousarray(fine_h.slab).copy()
    for st, dt in steps:
        orc.orc_sweep_multires(fine_h.sx, fine_h.sy, fine_h.sz, 4, slab, fine_h.b2s, coarse_h.slab, coarse_h.b2s,
                               sc.levels, active, active.size, st, dt, 1, 0, 8)
    assert_bitexact(fine.to_host().slab, slab, "512^3 multires steps")


def test_sphere512_all_levels_graph_bitexact(pkg, orc, hip, sphere):
    """Every block of the 512^3 grid swept at its own level (32,768 blocks over
    4 levels) by one plan, MC19 / BiLap25 graphs alternating, against the
    oracle's all-level sweeps."""
    sc = sphere
    dev = [pkg.DeviceGrid.from_host(g, lib=hip) for g in sc.grids]
    dl = hip.to_device(sc.levels)
    active = np.arange(sc.levels.size, dtype=np.int64)
    say("512^3 all-level sweeps over", active.size, "blocks")
    plan = pkg.SweepPlan(dev[0], active, levels=dl, level_grids=dev)
    steps = [(1, 0.02), (2, 0.005)] * 2
    for st, dt in steps:
        plan.run_graph_async(st, dt, 1)
    hip.synchronize()
    slabs = [np.ascontiguousarray(g.slab, np.float32).copy() for g in sc.grids]
    dims = np.array([[g.sx, g.sy, g.sz] for g in sc.grids], np.int64).ravel()
    sl = (C.c_void_p * 4)(*[a.ctypes.data for a in slabs])
    b2 = (C.c_void_p * 4)(*[np.ascontiguousarray(g.b2s).ctypes.data for g in sc.grids])
    for st, dt in steps:
        orc.orc_sweep_levels(dims, sl, b2, sc.levels, 4, active, active.size, st, dt, 1, 0, 8)
    for l in range(4):
        assert_bitexact(dev[l].to_host().slab, slabs[l], "level %d" % l)


@pytest.fixture(scope="module")
def band_particles(sphere):
    from msbg_rs_amd import scenes
    return scenes.band_particles(sphere, 16 << 20)


def test_sphere512_morton_sort_16m_bitexact(pkg, orc, hip, band_particles):
    """Keys and the stable sort permutation of the 16M band particles."""
    from msbg_rs_amd import particles as P
    pos, _ = band_particles
    n = pos.shape[0]
    say("16M Morton sort", n)
    keys, perm = P.morton_sort(pos, lib=hip)
    k = np.zeros(n, np.uint32)
    orc.orc_morton_keys(pos, n, k)
    want = np.zeros(n, np.uint32)
    orc.orc_stable_sort_perm(k, n, want)
    assert np.array_equal(perm.get(), want)
    assert np.array_equal(keys.get(), k[want])


def test_sphere512_bspline_transfers_16m_sampled_bitexact(pkg, orc, hip, band_particles):
    """Quadratic B-spline P2G of the 16M band particles over the 512^3 grid.
    The oracle runs on the particles of the 27 blocks around each of 96
    sampled active blocks: a cell sums its 27 voxels' particles in stable
    Morton order, so those blocks' cells get the same terms in the same order
    and must match bit for bit.  G2P of 500k sampled particles from the
    device's P2G grid against the oracle's sampler."""
    from msbg_rs_amd import transfer as T
    pos, vel = band_particles
    n = pos.shape[0]
    s = (512, 512, 512)
    nb = 32
    say("16M B-spline transfers", n)
    tg = T.TransferGrid(*s, capacity_blocks=3 * 4288 + 4096, lib=hip)
    k = tg.p2g(2, pos, vel)
    m_dev = tg.mass.to_host()
    v_dev = [tg.velocity[a].to_host() for a in range(3)]
    act = np.nonzero(m_dev.b2s >= 2)[0]
    assert act.size == k
    rng = np.random.default_rng(12)
    pick = np.sort(rng.choice(act, size=96, replace=False))
    sub = np.nonzero(near_blocks(pos, pick, nb))[0]
    say("  oracle on", sub.size, "particles around", pick.size, "blocks")
    _, b2s, slabs = oracle_p2g_bspline(orc, 2, s, pos[sub], vel[sub])
    assert np.all(b2s[pick] >= 2)
    assert_bitexact(m_dev.slab[m_dev.b2s[pick]], slabs[0][b2s[pick]], "mass")
    for a in range(3):
        assert_bitexact(v_dev[a].slab[v_dev[a].b2s[pick]], slabs[1 + a][b2s[pick]], "velocity %d" % a)
    # G2P of a particle sample, from the device grid
    sel = np.sort(rng.choice(n, size=500_000, replace=False))
    got = tg.g2p(2, pos[sel]).get()
    vs = np.stack([np.ascontiguousarray(v.slab) for v in v_dev])
    assert all(np.array_equal(v.b2s, m_dev.b2s) for v in v_dev)
    want = oracle_g2p_bspline(orc, 2, s, m_dev.b2s, vs, pos[sel])
    assert_bitexact(got, want, "g2p")


# ---------------------------------------------------------------------------
# 1024^3 dam break, 64M particles
def test_dam_break_1024_flip_step_properties(pkg, orc, hip):
    """One FLIP step at the benchmarked size (FP16 band storage, multigrid, to
    1e-6 or the FP32 floor).
    - MAC P2G: on 48 sampled blocks every face mass and velocity is within
      1e-6 of the float64 sums of all particles that reach them.
    - The post-projection divergence equals the solve's residual b - A p on
      the band's cells up to FP32 rounding of the faces (a few 1e-6 of
      ||div||), and is >= 1e5 times smaller than before.
    - G2P: the MAC samples of u and u* and the FLIP/PIC blend of 200k sampled
      particles equal the oracle's on the device's grids, bit for bit."""
    from msbg_rs_amd import scenes
    from msbg_rs_amd.flip import FlipStep
    from msbg_rs_amd.particles import active_blocks
    from test_gpu_transfer import MAC_TOL
    S, n = 1024, 64 << 20
    say("dam break", S, n)
    pos, vel, mass = scenes.dam_break(S, n)
    dp, dv, dm = hip.to_device(pos), hip.to_device(vel), hip.to_device(mass)
    cap = int(active_blocks(S, S, S, dp, 3.0, lib=hip).size * 1.1) + 64
    fs = FlipStep(S, S, S, capacity_blocks=cap, lib=hip)
    ps, out, st = fs.run(dp, dv, dm, r_p=2.0, alpha=0.95, storage=1, tol=1e-6, multigrid=True, omega=1.0,
                         sweeps_per_check=2, max_sweeps=400)
    sv = st["solve"]
    say("  step %.1f ms, solve rel %.3g after %d sweeps" % (st["ms"]["total"], sv["rel_residual"], sv["sweeps"]))
    assert sv["rel_residual"] <= 2e-6
    layout, band, p, b, active, band_blocks = fs.last
    # divergence after projection vs the residual
    post = pkg.DeviceGrid(S, S, S, 4, 0.0, 0.0, capacity_blocks=active.size, lib=hip)
    post.copy_from(layout)
    fs.mac.divergence(post, active)
    r = band.field(0)
    band.residual(0, p, b, r)
    rows = layout.block_slots()[band_blocks]
    div_post = post.read_blocks(band_blocks).astype(np.float64)
    res = r.get().reshape(-1, 4096)[rows].astype(np.float64)
    div_pre = b.get().reshape(-1, 4096)[rows].astype(np.float64)
    n_pre, n_post = np.linalg.norm(div_pre), np.linalg.norm(div_post)
    n_diff = np.linalg.norm(div_post - res)
    say("  ||div|| %.4g -> %.4g, ||div' - r|| %.3g" % (n_pre, n_post, n_diff))
    assert n_post <= 1e-5 * n_pre
    assert n_diff <= 5e-6 * n_pre
    # MAC P2G on sampled blocks against the float64 oracle
    rng = np.random.default_rng(5)
    mac_act = fs.mac.velocity[0].value_blocks()
    pick = np.sort(rng.choice(mac_act, size=48, replace=False))
    sub = np.nonzero(near_blocks(pos, pick, S // 16))[0]
    say("  MAC oracle on", sub.size, "particles around", pick.size, "blocks")
    _, b2s, sums = oracle_p2g_mac_f64(orc, (S, S, S), pos[sub], vel[sub], mass[sub], 2.0)
    _, _, absu = oracle_p2g_mac_f64(orc, (S, S, S), pos[sub], np.abs(vel[sub]), mass[sub], 2.0)
    assert np.all(b2s[pick] >= 2)
    for a in range(3):
        M = fs.mac.mass[a].read_blocks(pick).astype(np.float64)
        V = fs.mac_tilde.velocity[a].read_blocks(pick).astype(np.float64)  # u* = the P2G's faces
        M64, P64, Pa = sums[3 + a][b2s[pick]], sums[a][b2s[pick]], absu[a][b2s[pick]]
        assert np.all(np.abs(M - M64) <= MAC_TOL * M64), a
        pm = M64 > 0
        V64 = np.where(pm, P64 / np.where(pm, M64, 1.0), 0.0)
        bound = MAC_TOL * np.where(pm, Pa / np.where(pm, M64, 1.0) + np.abs(V64), 0.0)
        assert np.all(np.abs(V - V64) <= bound), a
    # staggered G2P + blend of sampled particles, from the device's grids
    sel = np.sort(rng.choice(n, size=200_000, replace=False))
    pos_s = ps.get()[sel]
    _, vs, _ = fs.sorted
    uo = vs.get()[sel]
    got = out.get()[sel]
    u_h = [fs.mac.velocity[a].to_host() for a in range(3)]
    t_h = [fs.mac_tilde.velocity[a].to_host() for a in range(3)]
    assert all(np.array_equal(h.b2s, u_h[0].b2s) for h in u_h + t_h)
    ns = max(h.n_slots for h in u_h + t_h)
    su = np.ascontiguousarray(np.stack([h.slab[:ns] for h in u_h]))
    stl = np.ascontiguousarray(np.stack([h.slab[:ns] for h in t_h]))
    m = sel.size
    us, uts, blend = (np.zeros((m, 3), np.float32) for _ in range(3))
    clamp = int(pkg.BoundaryCondition.Clamp)
    orc.orc_mac_sample(S, S, S, su, ns, u_h[0].b2s, np.ascontiguousarray(pos_s), m, 0, clamp, us)
    orc.orc_mac_sample(S, S, S, stl, ns, u_h[0].b2s, np.ascontiguousarray(pos_s), m, 0, clamp, uts)
    orc.orc_g2p_blend(np.ascontiguousarray(uo), us, uts, m, 0.95, blend)
    assert_bitexact(got, blend, "new particle velocities")
