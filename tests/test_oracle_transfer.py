"""Pins the oracle's transfers: the reference's P2G/G2P unit tests
(fluid/p2g.rs:527-619, fluid/g2p.rs:62-129) and closed forms for the
B-spline transfers that have no reference counterpart."""
import numpy as np
import pytest

from transfer_helpers import (dense_from_slabs, oracle_g2p_bspline, oracle_p2g_bspline, oracle_p2g_mac,
                              oracle_p2g_mac_f64, scatter_bspline_f64)

S32 = (32, 32, 32)


def naive(orc, s, pos, vel, mass, r_p):
    sx, sy, sz = s
    d = np.zeros((6, sz, sy, sx), np.float32)
    pos = np.ascontiguousarray(pos, np.float32)
    orc.orc_p2g_mac_naive(sx, sy, sz, pos, np.ascontiguousarray(vel, np.float32),
                          np.ascontiguousarray(mass, np.float32), pos.shape[0], r_p, d)
    return d


def staged_dense(orc, s, pos, vel, mass, r_p, ref_commit_width=0):
    active, b2s, slabs = oracle_p2g_mac(orc, s, pos, vel, mass, r_p, ref_commit_width)
    return np.stack([dense_from_slabs(b2s, slabs[c], s) for c in range(6)])


@pytest.mark.parametrize("pos,vel", [([[8.0, 8.0, 8.0]], [[1.0, 2.0, 3.0]]),     # p2g_01
                                     ([[15.9, 8.0, 8.0]], [[1.0, 0.0, 0.0]]),    # p2g_03
                                     ([[0.1, 0.1, 0.1]], [[1.0, 1.0, 1.0]])])    # p2g_04
def test_staged_p2g_matches_naive_reference_tests(orc, pos, vel):
    a = staged_dense(orc, S32, pos, vel, [1.0], 2.0)
    b = naive(orc, S32, pos, vel, [1.0], 2.0)
    np.testing.assert_allclose(a, b, atol=1e-4, rtol=0)
    if pos[0][0] == 15.9:
        assert (a[3, 8, 8, 16:] > 0).any()  # mass reached the +x neighbour block


def test_staged_p2g_many_particles_matches_naive(orc):
    rng = np.random.default_rng(0)
    pos = rng.uniform(0.5, 47.5, size=(400, 3)).astype(np.float32)
    vel = rng.normal(size=(400, 3)).astype(np.float32)
    mass = rng.uniform(0.5, 2.0, size=400).astype(np.float32)
    s = (48, 48, 48)
    a = staged_dense(orc, s, pos, vel, mass, 2.0)
    b = naive(orc, s, pos, vel, mass, 2.0)
    np.testing.assert_allclose(a[3:], b[3:], rtol=2e-6, atol=1e-6)
    np.testing.assert_allclose(a[:3], b[:3], rtol=1e-4, atol=1e-4)


def dense_f64_brute_force(s, pos, vel, mass, r_p):
    """Eq. 6 splat as a dense per-particle scatter: every face cell of the
    domain within the particle's footprint box gets the float32 products
    w * m and w * (m u_a), summed in float64 (no windows, regions or commit).
    The weight is the staged splat's float32 expression (fluid/p2g.rs:300-316):
    dx = x - (p - 1/2), |d|^2 = dx^2 + (dz^2 + dy^2), df2 = |d|^2 + d_a + 1/4."""
    sx, sy, sz = s
    f = np.float32
    out = np.zeros((6, sz, sy, sx))
    r = f(r_p)
    inv_r2 = f(1.0) / (r * r)
    for p, u, m in zip(np.asarray(pos, f), np.asarray(vel, f), np.asarray(mass, f)):
        p0 = p - f(0.5)
        for a in range(3):
            of = np.array([f(1.0) if a == k else f(0.5) for k in range(3)], f)
            lo = np.maximum(np.ceil(p - r - of).astype(int), 0)
            hi = np.minimum(np.floor(p + r - of).astype(int), np.array([sx, sy, sz]) - 1)
            if (hi < lo).any():
                continue
            ix, iy, iz = (np.arange(lo[k], hi[k] + 1) for k in range(3))
            Z, Y, X = np.meshgrid(iz, iy, ix, indexing="ij")
            dx, dy, dz = X.astype(f) - p0[0], Y.astype(f) - p0[1], Z.astype(f) - p0[2]
            d2 = dx * dx + (dz * dz + dy * dy)
            df2 = d2 + (dx, dy, dz)[a] + f(0.25)
            w = np.maximum(f(1.0) - df2 * inv_r2, f(0.0))
            w = w * w * w
            sl = (slice(lo[2], hi[2] + 1), slice(lo[1], hi[1] + 1), slice(lo[0], hi[0] + 1))
            out[3 + a][sl] += (w * m).astype(np.float64)
            out[a][sl] += (w * (m * u[a])).astype(np.float64)
    return out


@pytest.mark.parametrize("r_p", [1.5, 2.0, 3.0])
def test_f64_splat_matches_dense_brute_force(orc, r_p):
    """The float64-accumulated staged splat (windows, colour order, 27-region
    commit) equals a dense per-particle scatter of the same products to
    float64 rounding, on a whole-block domain where the commit drops only
    cells outside the domain; a dropped region or a misplaced cell would
    differ by a whole contribution."""
    rng = np.random.default_rng(int(r_p * 10))
    s = (48, 32, 48)
    n = 300
    pos = rng.uniform(0.0, np.array(s) - 1e-3, size=(n, 3)).astype(np.float32)
    pos[:3] = [[15.9, 8.0, 8.0], [16.1, 15.95, 31.99], [0.05, 31.9, 47.9]]  # block faces, domain corners
    vel = rng.normal(size=(n, 3)).astype(np.float32)
    mass = rng.uniform(0.5, 2.0, n).astype(np.float32)
    _, b2s, sums = oracle_p2g_mac_f64(orc, s, pos, vel, mass, r_p)
    got = np.stack([dense_from_slabs(b2s, sums[c], s) for c in range(6)])
    want = dense_f64_brute_force(s, pos, vel, mass, r_p)
    scale = dense_f64_brute_force(s, pos, np.abs(vel), mass, r_p)
    assert np.all(np.abs(got - want) <= 1e-13 * scale + 1e-300)
    assert (want[3:] > 0).sum() > 10000


def test_f64_splat_within_float_reordering_of_staged(orc):
    """The float32 staged splat (the reference's order) is within the
    reordering bound of its float64 sums: for k positive terms a float32 sum
    is within (k - 1) 2^-24 of the exact one."""
    rng = np.random.default_rng(5)
    s = (48, 48, 48)
    n = 3000
    pos = rng.uniform(0.0, 48 - 1e-3, size=(n, 3)).astype(np.float32)
    vel = rng.normal(size=(n, 3)).astype(np.float32)
    mass = rng.uniform(0.5, 2.0, n).astype(np.float32)
    _, b2s, f32s = oracle_p2g_mac(orc, s, pos, vel, mass, 2.0)
    _, b2s64, sums = oracle_p2g_mac_f64(orc, s, pos, vel, mass, 2.0)
    assert np.array_equal(b2s, b2s64)
    ones = oracle_p2g_mac_f64(orc, s, pos, vel, np.ones(n, np.float32), 2.0)[2]
    for a in range(3):
        M, M64 = f32s[3 + a], sums[3 + a]
        k = np.maximum((ones[3 + a] > 0) * 64.0, 1.0)  # <= 64 particles reach a cell here
        assert np.all(np.abs(M - M64) <= k * 2.0 ** -24 * M64 + 1e-30)
        assert np.abs(M - M64).max() > 0  # the float32 order does round


def test_reference_margin_commit_spill(orc):
    """The reference commits 4 lanes of a 2-wide -x margin region (p2g.rs:396-406):
    for a particle in a block with bx >= 1 near its -x face the staged splat then
    adds centre-block staging columns to the next row of the -x neighbour block,
    cells no particle reaches.  ref_commit_width = 1 reproduces it; the naive
    oracle (p2g.rs:445-502) and the exact-width commit do not."""
    pos, vel = [[16.6, 8.0, 8.0]], [[1.0, 0.0, 0.0]]
    quirk = staged_dense(orc, S32, pos, vel, [1.0], 2.0, ref_commit_width=1)
    exact = staged_dense(orc, S32, pos, vel, [1.0], 2.0, ref_commit_width=0)
    ref = naive(orc, S32, pos, vel, [1.0], 2.0)
    np.testing.assert_allclose(exact, ref, atol=1e-5, rtol=0)
    diff = np.argwhere(np.abs(quirk[3] - ref[3]) > 1e-6)
    assert diff.size > 0
    # every spilled cell is at x in {0, 1} of block bx = 0, one row (y + 1) above a touched row
    assert set(diff[:, 2]) <= {0, 1}


def test_rigid_translation_roundtrip(orc):
    """p2g_02: a rigidly moving cloud samples back to its velocity (Linear, Clamp)."""
    u = np.array([1.5, -2.0, 0.5], np.float32)
    pos = np.array([[8.0, 8.0, 8.0], [12.0, 9.0, 7.0], [9.0, 14.0, 10.0]], np.float32)
    active, b2s, slabs = oracle_p2g_mac(orc, S32, pos, np.tile(u, (3, 1)), np.ones(3, np.float32), 2.0)
    out = np.zeros((3, 3), np.float32)
    orc.orc_mac_sample(32, 32, 32, np.ascontiguousarray(slabs[:3]), slabs.shape[1], b2s, pos, 3, 0, 0, out)
    np.testing.assert_allclose(out, np.tile(u, (3, 1)), atol=1e-4)


def test_radius_cull_and_zero_mass_guard(orc):
    """p2g_05 / p2g_07."""
    d = staged_dense(orc, S32, [[8.0, 8.0, 8.0]], [[1.0, 0.0, 0.0]], [1.0], 2.0)
    assert d[3, 8, 8, 10] == 0.0 and d[3, 8, 8, 7] > 0.0
    assert d[0, 20, 20, 20] == 0.0 and np.isfinite(d[0, 8, 8, 7])


def test_overlap_sum_order_independent(orc):
    """p2g_06."""
    a = staged_dense(orc, S32, [[8.0, 8.0, 8.0], [9.0, 8.0, 8.0]], [[1.0, 0.0, 0.0]] * 2, [1.0, 1.0], 2.0)
    b = staged_dense(orc, S32, [[9.0, 8.0, 8.0], [8.0, 8.0, 8.0]], [[1.0, 0.0, 0.0]] * 2, [1.0, 1.0], 2.0)
    np.testing.assert_allclose(a, b, atol=1e-5)


def face_field_slabs(b2s_n=8):
    """g2p_11 grid: u(i,j,k) = i+1, v = j+1, w = k+1 over a 32^3 domain."""
    b2s = np.arange(2, 2 + b2s_n, dtype=np.int32)
    slabs = np.zeros((3, 2 + b2s_n, 4096), np.float32)
    v = np.arange(4096)
    lx, ly, lz = v & 15, (v >> 4) & 15, v >> 8
    for bid in range(8):
        bx, by, bz = bid % 2, (bid // 2) % 2, bid // 4
        slabs[0, 2 + bid] = bx * 16 + lx + 1
        slabs[1, 2 + bid] = by * 16 + ly + 1
        slabs[2, 2 + bid] = bz * 16 + lz + 1
    return b2s, slabs


def test_g2p_face_center_exact(orc):
    """g2p_11: sampling at a face centre returns that face value exactly."""
    b2s, slabs = face_field_slabs()
    out = np.zeros((1, 3), np.float32)
    orc.orc_mac_sample(32, 32, 32, slabs, slabs.shape[1], b2s, np.array([[5.0, 8.5, 8.5]], np.float32), 1, 0, 0, out)
    assert list(out[0]) == [5.0, 8.5, 8.5]


def test_g2p_blend_endpoints(orc):
    """g2p_08: alpha 0 = PIC (u), alpha 1 = FLIP (u_old + u - u_tilde)."""
    uo, u, ut = (np.array([[1, 2, 3]], np.float32), np.array([[4, 5, 6]], np.float32),
                 np.array([[4.5, 4.5, 4.5]], np.float32))
    out = np.zeros((1, 3), np.float32)
    orc.orc_g2p_blend(uo, u, ut, 1, 0.0, out)
    assert np.array_equal(out, u)
    orc.orc_g2p_blend(uo, u, ut, 1, 1.0, out)
    assert np.array_equal(out, uo + (u - ut))


@pytest.mark.parametrize("order", [1, 2])
def test_bspline_partition_of_unity_and_closed_form(orc, order):
    rng = np.random.default_rng(order)
    for p in rng.uniform(2, 30, 200).astype(np.float32):
        f = int(np.floor(p))
        w = [orc.orc_bspline_weight(order, float(p), c) for c in range(f - 2, f + 3)]
        assert abs(sum(w) - 1.0) < 2e-7
        x = float(np.float32(p) - np.float32(0.5)) - f  # distance to cell f's centre
        if order == 2:
            assert abs(w[2] - (0.75 - x * x)) < 1e-6  # quadratic B-spline 3/4 - x^2
        else:
            assert abs(w[2] - max(0.0, 1.0 - abs(x))) < 1e-6
    assert orc.orc_bspline_weight(2, 10.5, 10) == 0.75 and orc.orc_bspline_weight(2, 10.5, 11) == 0.125
    assert orc.orc_bspline_weight(1, 10.5, 10) == 1.0 and orc.orc_bspline_weight(1, 10.5, 11) == 0.0


@pytest.mark.parametrize("order", [1, 2])
def test_bspline_p2g_conservation_and_scatter_equivalence(orc, order):
    rng = np.random.default_rng(10 + order)
    s = (48, 40, 32)
    pos = rng.uniform(2.0, [46.0, 38.0, 30.0], size=(300, 3)).astype(np.float32)
    vel = rng.normal(size=(300, 3)).astype(np.float32)
    mass = rng.uniform(0.5, 2.0, 300).astype(np.float32)
    active, b2s, slabs = oracle_p2g_bspline(orc, order, s, pos, vel, mass)
    M = dense_from_slabs(b2s, slabs[0], s).astype(np.float64)
    V = np.stack([dense_from_slabs(b2s, slabs[1 + a], s) for a in range(3)]).astype(np.float64)
    assert abs(M.sum() - mass.astype(np.float64).sum()) < 1e-4  # mass conservation
    mom = (M[None] * V).reshape(3, -1).sum(1)
    np.testing.assert_allclose(mom, (mass[:, None] * vel).astype(np.float64).sum(0), atol=1e-3)  # momentum
    Mb, Pb = scatter_bspline_f64(orc, order, s, pos, vel, mass)
    np.testing.assert_allclose(M, Mb, rtol=1e-5, atol=1e-6)
    Vb = np.where(Mb > 0, Pb / np.where(Mb > 0, Mb, 1), 0)
    np.testing.assert_allclose(V, Vb, rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("order", [1, 2])
def test_bspline_g2p_reproduces_affine_field(orc, order):
    """B-splines of order 1 and 2 reproduce affine functions exactly."""
    s = (48, 48, 48)
    nb = 27
    b2s = np.arange(2, 2 + nb, dtype=np.int32)
    slabs = np.zeros((3, 2 + nb, 4096), np.float32)
    v = np.arange(4096)
    for bid in range(nb):
        bx, by, bz = bid % 3, (bid // 3) % 3, bid // 9
        cx, cy, cz = bx * 16 + (v & 15) + 0.5, by * 16 + ((v >> 4) & 15) + 0.5, bz * 16 + (v >> 8) + 0.5
        slabs[0, 2 + bid] = 0.5 * cx - 0.25 * cy + 2.0
        slabs[1, 2 + bid] = 0.1 * cz + 1.0
        slabs[2, 2 + bid] = -0.3 * cx + 0.2 * cy + 0.1 * cz
    rng = np.random.default_rng(order)
    pos = rng.uniform(3, 45, size=(500, 3)).astype(np.float32)
    out = oracle_g2p_bspline(orc, order, s, b2s, slabs, pos)
    p = pos.astype(np.float64)
    want = np.stack([0.5 * p[:, 0] - 0.25 * p[:, 1] + 2.0, 0.1 * p[:, 2] + 1.0,
                     -0.3 * p[:, 0] + 0.2 * p[:, 1] + 0.1 * p[:, 2]], 1)
    np.testing.assert_allclose(out, want, atol=2e-5)


def mac_channels(pkg, s, f, frac=1.0, seed=0):
    """Three velocity channels (each its own grid) filled with f(axis, X, Y, Z)."""
    rng = np.random.default_rng(seed)
    out = []
    for a in range(3):
        g = pkg.SparseGrid("u%d" % a, *s, 0.0, 0.0)
        bids = np.nonzero(rng.random(g.n_blocks) < frac)[0] if frac < 1 else np.arange(g.n_blocks)
        g.ensure_blocks_parallel(bids)
        g.fill_from_function(bids, lambda X, Y, Z, a=a: f(a, X, Y, Z).astype(np.float32))
        out.append(g)
    return out


def oracle_divergence(pkg, orc, chans, active, ref_packed=0):
    s = (chans[0].sx, chans[0].sy, chans[0].sz)
    d = pkg.SparseGrid("div", *s, 0.0, 0.0)
    d.ensure_blocks_parallel(active)
    slab = np.zeros((d.n_slots, 4096), np.float32)
    u, v, w = (np.ascontiguousarray(c.slab) for c in chans)
    orc.orc_mac_divergence(*s, u, chans[0].b2s, v, chans[1].b2s, w, chans[2].b2s, slab, d.b2s,
                           np.ascontiguousarray(active, np.int64), len(active), ref_packed)
    d._slab[: d.n_slots] = slab
    return d


def test_div_10_uniform_and_linear(pkg, orc):
    """fluid/mac.rs:193-225: u(i) = i + 1, v(j) = j + 1, w(k) = k + 1 -> 3 everywhere
    (the index-0 faces read 0)."""
    chans = mac_channels(pkg, (32, 32, 32), lambda a, X, Y, Z: (X, Y, Z)[a] + 1.0)
    d = oracle_divergence(pkg, orc, chans, np.arange(8))
    assert np.all(d.to_dense() == 3.0)


def test_divergence_layout_index_and_reference_packing(pkg, orc):
    """Whole blocks: the reference's running index is the layout index.  A
    partial block (domain 40 = 2.5 blocks): the reference packs the clipped
    extent, the layout index puts every cell where get_voxel reads it."""
    f = lambda a, X, Y, Z: np.sin(0.3 * X + 0.2 * a) * np.cos(0.1 * Y) + 0.05 * Z * (a + 1)
    chans = mac_channels(pkg, (40, 32, 32), f)
    act = np.arange(chans[0].n_blocks)
    exact = oracle_divergence(pkg, orc, chans, act, 0)
    packed = oracle_divergence(pkg, orc, chans, act, 1)
    D = exact.to_dense()
    X = [c.to_dense().astype(np.float64) for c in chans]
    want = np.zeros_like(D, np.float64)
    for a, ax in ((0, 2), (1, 1), (2, 0)):
        lo = np.concatenate([np.zeros_like(np.take(X[a], [0], axis=ax)), np.take(X[a], range(X[a].shape[ax] - 1),
                                                                                  axis=ax)], axis=ax)
        want = want + (X[a] - lo)
    assert np.allclose(D, want, rtol=0, atol=1e-5)
    whole = [b for b in act if (b % 3) < 2]  # blocks 0, 1 along x are whole, block 2 holds x 32..39
    assert np.array_equal(exact.slab[exact.b2s[whole]], packed.slab[packed.b2s[whole]])
    part = [b for b in act if (b % 3) == 2]
    assert not np.array_equal(exact.slab[exact.b2s[part]], packed.slab[packed.b2s[part]])
