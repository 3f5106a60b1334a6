"""Pins of the oracle's multires transfers with the reference's own tests:
downsample_channel_avg (multires/downsample.rs:97-209, test_ds_01..04) and
regularize_refinement_map (multires/refinement.rs:532-629, test_ref_01/03/05/06),
plus the 2:1 invariant on random maps."""
import numpy as np


def fine_and_coarse(pkg, sx, sy, sz, f):
    """Level 0 (16^3 blocks) filled with f(x, y, z) on every block, and an
    empty level 1 (8^3 blocks) over the same block grid."""
    fine = pkg.SparseGrid("fine", sx, sy, sz, 0.0, 0.0)
    bids = np.arange(fine.n_blocks, dtype=np.int64)
    fine.ensure_blocks_parallel(bids)
    fine.fill_from_function(bids, lambda X, Y, Z: f(X, Y, Z).astype(np.float32))
    coarse = pkg.SparseGrid("coarse", sx // 2, sy // 2, sz // 2, 0.0, 0.0, log2_bsx=3)
    coarse.ensure_blocks_parallel(bids)
    return fine, coarse, bids


def downsample(orc, fine, coarse, bids):
    cs = np.ascontiguousarray(coarse.slab).copy()
    orc.orc_downsample_avg(4, np.ascontiguousarray(fine.slab), fine.b2s, cs, coarse.b2s, bids, bids.size)
    coarse._slab[: coarse.n_slots] = cs
    return coarse.to_dense()


def test_ds_01_constant_field_exact(pkg, orc):
    fine, coarse, bids = fine_and_coarse(pkg, 32, 32, 32, lambda X, Y, Z: np.full(X.shape, 0.5))
    assert np.all(downsample(orc, fine, coarse, bids) == 0.5)


def test_ds_02_linear_gradient(pkg, orc):
    fine, coarse, bids = fine_and_coarse(pkg, 32, 16, 16, lambda X, Y, Z: X.astype(np.float64))
    got = downsample(orc, fine, coarse, bids)
    assert np.all(got == (2.0 * np.arange(16) + 0.5)[None, None, :])


def test_ds_03_bilinear_closed_form(pkg, orc):
    """f = x y + z: the 2x2x2 mean is (2cx + 1/2)(2cy + 1/2) + 2cz + 1/2, exact in FP32."""
    fine, coarse, bids = fine_and_coarse(pkg, 32, 16, 16, lambda X, Y, Z: (X * Y + Z).astype(np.float64))
    got = downsample(orc, fine, coarse, bids)
    cz, cy, cx = np.meshgrid(np.arange(8), np.arange(8), np.arange(16), indexing="ij")
    assert np.all(got == (2 * cx + 0.5) * (2 * cy + 0.5) + 2 * cz + 0.5)
    # sample_coarse at (2cx, 2cy, 2cz) is that coarse voxel
    cs = np.ascontiguousarray(coarse.slab)
    for x, y, z in [(0, 0, 0), (15, 7, 7), (3, 5, 1)]:
        v = orc.orc_sample_coarse(16, 8, 8, 3, cs, coarse.b2s, 2 * x, 2 * y, 2 * z)
        assert v == got[z, y, x]


def test_ds_04_out_of_domain_clamps(pkg, orc):
    coarse = pkg.SparseGrid("coarse", 8, 8, 8, 0.0, 0.0, log2_bsx=3)
    coarse.ensure_blocks_parallel(np.array([0], np.int64))
    slab = np.ascontiguousarray(coarse.slab).copy()
    slab[coarse.b2s[0], 0] = 7.0
    slab[coarse.b2s[0], 511] = 9.0
    assert orc.orc_sample_coarse(8, 8, 8, 3, slab, coarse.b2s, -3, -3, -3) == 7.0
    assert orc.orc_sample_coarse(8, 8, 8, 3, slab, coarse.b2s, 100, 100, 100) == 9.0


def regularize(orc, m, n=(4, 4, 4), levels=3):
    m = np.ascontiguousarray(m, np.uint8).copy()
    orc.orc_regularize(m, *n, levels)
    return m


def test_ref_01_uniform_unchanged(orc):
    for v in (0, 2):
        assert np.all(regularize(orc, np.full(64, v)) == v)


def test_ref_03_island_gets_a_staircase(orc):
    """A level-0 block in a level-2 sea: exactly its 26 neighbours are lifted to
    level 1, everything else stays (and the 2:1 invariant holds)."""
    m = np.full(64, 2, np.uint8)
    m[21] = 0  # block (1, 1, 1)
    got = regularize(orc, m)
    want = np.full((4, 4, 4), 2, np.uint8)
    want[0:3, 0:3, 0:3] = 1
    want[1, 1, 1] = 0
    assert np.array_equal(got.reshape(4, 4, 4), want)


def test_ref_05_ref_06_small_level_counts_and_idempotence(orc):
    m = np.zeros(64, np.uint8)
    assert np.all(regularize(orc, m, levels=1) == 0) and np.all(regularize(orc, m, levels=2) == 0)
    once = regularize(orc, m)
    assert np.array_equal(regularize(orc, once), once)


def test_regularized_random_maps_are_2_to_1(orc):
    rng = np.random.default_rng(0)
    n = (9, 7, 8)
    for levels in (3, 4, 5):
        for _ in range(5):
            m = rng.integers(0, levels, size=n[0] * n[1] * n[2]).astype(np.uint8)
            got = regularize(orc, m, n, levels).reshape(n[2], n[1], n[0])
            assert np.all(got <= m.reshape(n[2], n[1], n[0]))  # levels only refine
            for dz in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        a = got[max(dz, 0):n[2] + min(dz, 0), max(dy, 0):n[1] + min(dy, 0),
                                max(dx, 0):n[0] + min(dx, 0)].astype(int)
                        b = got[max(-dz, 0):n[2] + min(-dz, 0), max(-dy, 0):n[1] + min(-dy, 0),
                                max(-dx, 0):n[0] + min(-dx, 0)].astype(int)
                        assert np.abs(a - b).max() <= 1
