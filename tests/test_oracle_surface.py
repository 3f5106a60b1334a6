"""Pins of the oracle's surface-reconstruction tail and block topology:

- finalize (particles/finalize.rs:35-71): the reference's finalize_01..05
  known answers (finalize.rs:107-183), and every u16 code against the
  float64 closed form;
- place (particles/sort.rs:31-143): a two-point cloud worked out by hand,
  and the domain bounds (particles/mod.rs:180-194) against their closed form;
- compute_block_topology (multires/refinement.rs:124-208): test_ref_02 and
  test_ref_04 (refinement.rs:548-613), and a vectorised numpy formulation of
  the 26-neighbour min / max on random 2:1 maps.
"""
import numpy as np
import pytest

R_P, NB = 2.0, 2.0  # finalize_0x's SurfaceConfig: r_particle 2, nb_dist 2 (finalize.rs:88-99)


def finalize_grid(pkg, s):
    g = pkg.SparseGrid("finalize", *s, 0.0, 1.0)
    g.ensure_blocks_parallel(np.arange(g.n_blocks))
    slab = np.zeros((g.n_slots, 4096), np.uint16)
    slab[1] = 65535
    return g, slab


def set_voxel(g, slab, x, y, z, u):
    slab[g.b2s[g.get_block_id(x, y, z)], (x & 15) | ((y & 15) << 4) | ((z & 15) << 8)] = u


def get_voxel(g, slab, x, y, z):
    return float(slab[g.b2s[g.get_block_id(x, y, z)], (x & 15) | ((y & 15) << 4) | ((z & 15) << 8)]) / 65535.0


def finalize(orc, g, slab, active, r_p=R_P, nb=NB):
    out = slab.copy()
    a = np.ascontiguousarray(active, np.int64)
    orc.lib.orc_density_finalize(g.b2s, out.reshape(-1), a, a.size, r_p, nb)
    return out


def test_finalize_01_single_particle(pkg, orc):
    g, slab = finalize_grid(pkg, (32, 32, 32))
    slab[2:] = 65535
    set_voxel(g, slab, 8, 8, 8, 0)
    set_voxel(g, slab, 6, 8, 8, round(2.5 * 2.5 / 16.0 * 65535.0))
    set_voxel(g, slab, 4, 8, 8, 65468)
    out = finalize(orc, g, slab, [0])
    assert abs(get_voxel(g, out, 8, 8, 8) - 1.0) < 1e-4
    assert abs(get_voxel(g, out, 6, 8, 8) - (1.0 - (0.5 + 2.0) / 6.0)) < 1e-4
    ratio = np.float32(65468.0) / np.float32(65535.0)
    assert abs(get_voxel(g, out, 4, 8, 8) - (1.0 - float(np.sqrt(ratio * 16.0)) / 6.0)) < 1e-4


def test_finalize_02_untouched_is_zero(pkg, orc):
    g, slab = finalize_grid(pkg, (32, 32, 32))
    set_voxel(g, slab, 3, 3, 3, 65535)
    out = finalize(orc, g, slab, [0])
    assert get_voxel(g, out, 3, 3, 3) == 0.0


def test_finalize_03_threshold_edge(pkg, orc):
    g, slab = finalize_grid(pkg, (32, 32, 32))
    set_voxel(g, slab, 0, 0, 0, round(0.9989 * 65535.0))
    set_voxel(g, slab, 1, 0, 0, round(0.9991 * 65535.0))
    out = finalize(orc, g, slab, [0])
    assert get_voxel(g, out, 0, 0, 0) > 0.0
    assert get_voxel(g, out, 1, 0, 0) == 0.0


def test_finalize_04_mixed_block_in_range(pkg, orc):
    g, slab = finalize_grid(pkg, (16, 16, 16))
    z, y, x = np.meshgrid(np.arange(16), np.arange(16), np.arange(16), indexing="ij")
    ratio = ((x + y + z) % 17) / 17.0
    slab[2] = np.round(ratio.reshape(-1) * 65535.0).astype(np.uint16)
    out = finalize(orc, g, slab, [0])
    v = out[2].astype(np.float64) / 65535.0
    assert np.all((v >= 0.0) & (v <= 1.0))


def test_finalize_05_empty_active(pkg, orc):
    g, slab = finalize_grid(pkg, (16, 16, 16))
    out = finalize(orc, g, slab, [])
    assert np.array_equal(out, slab)


@pytest.mark.parametrize("r_p,nb", [(2.0, 2.0), (1.5, 3.0)])
def test_finalize_every_code_against_closed_form(pkg, orc, r_p, nb):
    """All 65536 codes: density = 1 - clamp(sqrt(ratio) r_scan / (2 r_p + nb)),
    0 at ratio >= 0.999, evaluated in float64; the float32 pipeline may land
    one code away where the float64 value sits at a rounding boundary."""
    g, slab = finalize_grid(pkg, (64, 64, 16))
    slab[2:18] = np.arange(65536, dtype=np.uint32).astype(np.uint16).reshape(16, 4096)
    out = finalize(orc, g, slab, np.arange(16), r_p, nb)[2:18].reshape(-1).astype(np.int64)
    u = np.arange(65536, dtype=np.float64)
    ratio = u / 65535.0
    r_scan = r_p + nb
    f = 1.0 - np.clip(np.sqrt(ratio) * r_scan / (2.0 * r_p + nb), 0.0, 1.0)
    want = np.where(ratio < 0.999, np.round(f * 65535.0), 0.0).astype(np.int64)
    edge = np.abs(ratio - 0.999) < 1e-6  # the float32 threshold comparison may fall either way here
    assert np.all(np.abs(out - want)[~edge] <= 1)
    assert int(np.abs(out - want)[~edge].max()) <= 1 and int((out != want).sum()) < 200


def test_domain_bounds_closed_form(orc):
    out = np.zeros(6, np.float32)
    orc.lib.orc_domain_bounds(32, 48, 100, out)
    scale = 16777215.0 / 32768.0  # 24-bit fixed point over a 32768 range
    lo = np.floor(scale * 0.005 + 0.5) / scale
    assert out[0] == out[1] == out[2] == np.float32(lo)
    assert abs(lo - 3.0 / 511.99997) < 1e-9  # floor(2.56 + 0.5) = 3 steps of 1/scale
    for k, s in enumerate((32, 48, 100)):
        assert out[3 + k] == np.float32(np.floor(scale * (s - 0.005) + 0.5) / scale)
        assert s - 0.006 < out[3 + k] < s


def test_place_two_points_by_hand(orc):
    """Base points (0,0,0) and (1,1,1), bbox min 0, span 1, 32^3, factor 0.5:
    smax = smin = 32, inst_scale = 16; instance origins 6.4 and 6.4 + 19.2 =
    25.6; particles 6.4, 22.4, 25.6 and 41.6 (outside, dropped)."""
    base = np.array([[0, 0, 0], [1, 1, 1]], np.float32)
    pos = np.zeros((8, 3), np.float32)
    bids = np.zeros(8, np.int64)
    k = orc.lib.orc_surface_place(base.reshape(-1), 2, np.zeros(3, np.float32), 1.0, 32, 32, 32, 0, 0.5,
                                  pos.reshape(-1), bids, 8)
    assert k == 3
    np.testing.assert_allclose(pos[:3, 0], [6.4, 22.4, 25.6], rtol=1e-6)
    assert np.all(pos[:3, 0] == pos[:3, 1]) and np.all(pos[:3, 0] == pos[:3, 2])
    assert list(bids[:3]) == [0, 7, 7]  # 2x2x2 blocks: (0,0,0) -> 0, (1,1,1) -> 1 + 2 + 4
    assert orc.lib.orc_surface_place(base.reshape(-1), 2, np.zeros(3, np.float32), 1.0, 32, 32, 32, 0, 0.5,
                                     pos.reshape(-1), bids, 2) == -1  # capacity


def test_place_instances_and_domain(orc):
    """Every kept particle lies inside the bounds and the grid; each instance
    is a scaled, translated copy of the cloud (origin from its own point)."""
    rng = np.random.default_rng(4)
    base = rng.normal(size=(300, 3)).astype(np.float32)
    bmin = base.min(0)
    span = float((base.max(0) - bmin).max())
    n_inst = 40
    cap = n_inst * 300
    pos = np.zeros((cap, 3), np.float32)
    bids = np.zeros(cap, np.int64)
    s = (64, 48, 80)
    k = orc.lib.orc_surface_place(base.reshape(-1), 300, bmin, span, *s, n_inst, 0.05, pos.reshape(-1), bids, cap)
    assert 0 < k < cap
    b = np.zeros(6, np.float32)
    orc.lib.orc_domain_bounds(*s, b)
    p = pos[:k]
    assert np.all(p >= b[:3]) and np.all(p <= b[3:])
    ip = p.astype(np.int64)
    assert np.array_equal(bids[:k], (ip[:, 0] >> 4) + (ip[:, 1] >> 4) * 4 + (ip[:, 2] >> 4) * 12)
    # float64 positions of all n_inst x 300 candidates: the kept ones, in order
    smax, smin = 80.0, 48.0
    origin = 0.2 * smax + 0.6 * smax / span * (base[:n_inst].astype(np.float64) - bmin)
    cand = origin[:, None, :] + smin * 0.05 / span * (base[None].astype(np.float64) - bmin)
    inside = np.all((cand >= b[:3]) & (cand <= b[3:]), axis=2).reshape(-1)
    near = np.any((np.abs(cand - b[:3]) < 1e-4) | (np.abs(cand - b[3:]) < 1e-4), axis=2).reshape(-1)
    assert not near.any()  # no candidate within rounding of a bound for this cloud
    assert k == int(inside.sum())
    np.testing.assert_allclose(p, cand.reshape(-1, 3)[inside], atol=1e-4)


FC, CF, DOM, EXISTS = 1 << 3, 1 << 2, 1 << 1, 1 << 12


def topology(orc, levels, dims, n_levels, user=None):
    nx, ny, nz = dims
    flags = np.zeros(levels.size, np.uint16)
    fc = np.zeros(levels.size, np.int64)
    counts = np.zeros(n_levels, np.int64)
    u = None if user is None else np.ascontiguousarray(user, np.uint16)
    n_act = orc.lib.orc_block_topology(nx, ny, nz, np.ascontiguousarray(levels, np.uint8), n_levels,
                                       None if u is None else u.ctypes.data, flags, fc, counts)
    lists, k = [], 0
    for c in counts:
        lists.append(list(fc[k:k + c]))
        k += c
    return flags, lists, n_act


def test_ref_02_single_fine_block(orc):
    levels = np.ones(64, np.uint8)
    levels[21] = 0
    flags, lists, _ = topology(orc, levels, (4, 4, 4), 3)
    assert flags[21] & FC and not flags[21] & CF
    assert lists[0] == [21]
    assert not flags[63] & (FC | CF)


def test_ref_04_dom_border(orc):
    flags, _, _ = topology(orc, np.zeros(64, np.uint8), (4, 4, 4), 3)
    assert flags[0] & DOM and flags[63] & DOM
    assert not flags[21] & DOM


def test_topology_matches_numpy_formulation(orc):
    rng = np.random.default_rng(5)
    dims = (9, 7, 6)
    nx, ny, nz = dims
    lv = rng.integers(0, 3, size=nx * ny * nz).astype(np.uint8)
    orc.orc_regularize(lv, nx, ny, nz, 3)  # a legal 2:1 map, as the reference's callers pass
    user = (rng.random(lv.size) < 0.5).astype(np.uint16) * EXISTS | (1 << 5)
    flags, lists, n_act = topology(orc, lv, dims, 3, user)
    L = lv.reshape(nz, ny, nx).astype(np.int64)
    hi = np.pad(L, 1, constant_values=-1)
    lo = np.pad(L, 1, constant_values=99)
    mx = np.full_like(L, -1)
    mn = np.full_like(L, 99)
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                if (dx, dy, dz) == (1, 1, 1):
                    continue
                mx = np.maximum(mx, hi[dz:dz + nz, dy:dy + ny, dx:dx + nx])
                mn = np.minimum(mn, lo[dz:dz + nz, dy:dy + ny, dx:dx + nx])
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    border = (x == 0) | (x == nx - 1) | (y == 0) | (y == ny - 1) | (z == 0) | (z == nz - 1)
    want = user.reshape(nz, ny, nx).astype(np.int64) | np.where(border, DOM, 0) | np.where(mx > L, FC, 0) | \
        np.where(mn < L, CF, 0)
    assert np.array_equal(flags.reshape(nz, ny, nx).astype(np.int64), want)
    for l in range(3):
        assert lists[l] == list(np.nonzero(((want & FC) != 0).reshape(-1) & (lv == l))[0])
    assert n_act == int(((user & EXISTS) != 0).sum())
