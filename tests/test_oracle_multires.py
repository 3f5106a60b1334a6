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


# ---------------------------------------------------------------------------
# Halos and sweeps of blocks at every level (orc_halo_fill_level,
# orc_sweep_levels).  No reference counterpart: fill_multires only fills a
# fine centre block.  Pinned by the independently pinned pieces each halo
# voxel must equal -- the single-level halo fill (golden-pinned), sample_coarse
# and downsample_channel_avg (test_ds_*) -- by constant fields staying
# constant across levels, and by reducing to the fine-only multires sweep.
# ---------------------------------------------------------------------------
import ctypes as C  # noqa: E402

import pytest  # noqa: E402


class LevelGrid:
    """A multires grid for the oracle: per level l a SparseGrid with block edge
    16 >> l over one block grid, holding the blocks whose level is l."""

    def __init__(self, pkg, orc, s, n_levels, seed, lo=-1.0, hi=1.0, const=None, frac_l0=0.08, frac_mixed=0.5):
        rng = np.random.default_rng(seed)
        nb = [(v + 15) // 16 for v in s]
        n = nb[0] * nb[1] * nb[2]
        m = np.full(n, n_levels - 1, np.uint8)
        u = rng.random(n)
        m[u < frac_mixed] = rng.integers(1, n_levels, size=int((u < frac_mixed).sum()))
        m[u < frac_l0] = 0
        orc.orc_regularize(m, nb[0], nb[1], nb[2], n_levels)
        self.s, self.nb, self.n_levels, self.levels = s, nb, n_levels, m
        self.grids = []
        for l in range(n_levels):
            g = pkg.SparseGrid("l%d" % l, s[0] >> l, s[1] >> l, s[2] >> l, 0.0, 0.0, log2_bsx=4 - l)
            bids = np.nonzero(m == l)[0].astype(np.int64)
            g.ensure_blocks_parallel(bids)
            if const is None:
                g._slab[2:g.n_slots] = rng.uniform(lo, hi, size=(g.n_slots - 2, g.N)).astype(np.float32)
            else:
                g._slab[2:g.n_slots] = np.float32(const)
            self.grids.append(g)
        self.slabs = [np.ascontiguousarray(g.slab, np.float32).copy() for g in self.grids]

    def dims(self):
        return np.array([[g.sx, g.sy, g.sz] for g in self.grids], np.int64).ravel()

    def ptrs(self):
        sl = (C.c_void_p * self.n_levels)(*[a.ctypes.data for a in self.slabs])
        b2 = (C.c_void_p * self.n_levels)(*[np.ascontiguousarray(g.b2s).ctypes.data for g in self.grids])
        return sl, b2

    def fill(self, orc, l, bid, halo):
        t = 16 >> l
        tile = np.zeros((t + 2 * halo) ** 3, np.float32)
        sl, b2 = self.ptrs()
        orc.orc_halo_fill_level(self.dims(), sl, b2, self.levels, self.n_levels, l, int(bid), halo, tile)
        return tile.reshape((t + 2 * halo,) * 3)

    def sweep(self, orc, active, stencil, dt, iterations, clamp01=False, nthreads=4):
        a = np.ascontiguousarray(active, np.int64)
        sl, b2 = self.ptrs()
        orc.orc_sweep_levels(self.dims(), sl, b2, self.levels, self.n_levels, a, a.size, stencil, dt, iterations,
                             int(clamp01), nthreads)


def _reflect(g, mx):
    period = 2 * (mx + 1)
    r = g % period
    return np.where(r > mx, period - 1 - r, r)


@pytest.mark.parametrize("halo", [1, 2])
@pytest.mark.parametrize("n_levels", [3, 4])
def test_level_fill_equals_pinned_pieces(pkg, orc, halo, n_levels):
    """Every halo voxel of every coarse block equals: the single-level halo
    fill of its own level where its block lives there; sample_coarse of the
    next coarser level where its block is coarser; downsample_channel_avg of
    the finer neighbour where its block is finer."""
    lg = LevelGrid(pkg, orc, (160, 128, 128), n_levels, seed=halo + 10 * n_levels)
    nbx, nby, _ = lg.nb
    down = {}  # level l -> level-(l-1) data downsampled into level l's layout
    for l in range(1, n_levels):
        fine, dst = lg.grids[l - 1], lg.grids[l]
        bids = np.nonzero(lg.levels == l - 1)[0].astype(np.int64)
        tmp = pkg.SparseGrid("d", dst.sx, dst.sy, dst.sz, 0.0, 0.0, log2_bsx=dst.L)
        tmp.ensure_blocks_parallel(bids)
        cs = np.ascontiguousarray(tmp.slab).copy()
        orc.orc_downsample_avg(fine.L, lg.slabs[l - 1], fine.b2s, cs, tmp.b2s, bids, bids.size)
        down[l] = (tmp, cs)
    checked = {0: 0, 1: 0, 2: 0}
    for l in range(1, n_levels):
        g = lg.grids[l]
        E, L = 16 >> l, 4 - l
        cand = np.nonzero(lg.levels == l)[0]
        lv3 = lg.levels.reshape(lg.nb[2], lg.nb[1], lg.nb[0]).astype(int)
        pad = np.pad(lv3, 1, mode="edge")
        nmax = np.max([pad[1 + dz:pad.shape[0] - 1 + dz, 1 + dy:pad.shape[1] - 1 + dy, 1 + dx:pad.shape[2] - 1 + dx]
                       for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)], axis=0).ravel()
        nmin = np.min([pad[1 + dz:pad.shape[0] - 1 + dz, 1 + dy:pad.shape[1] - 1 + dy, 1 + dx:pad.shape[2] - 1 + dx]
                       for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)], axis=0).ravel()
        mixed = cand[(nmax[cand] > l) | (nmin[cand] < l)]
        for bid in np.concatenate([mixed[:30], cand[:5]]):
            tile = lg.fill(orc, l, bid, halo)
            own = orc.halo_fill(_as_grid(g, lg.slabs[l]), int(bid), halo, True, 1).reshape(tile.shape)
            bx, by, bz = bid % nbx, (bid // nbx) % nby, bid // (nbx * nby)
            h = np.arange(E + 2 * halo)
            rz, ry, rx = np.meshgrid(_reflect(bz * E + h - halo, g.sz - 1), _reflect(by * E + h - halo, g.sy - 1),
                                     _reflect(bx * E + h - halo, g.sx - 1), indexing="ij")
            nbl = lg.levels[(rx >> L) + (ry >> L) * nbx + (rz >> L) * nbx * nby]
            same = (nbl == l) | ((nbl > l) & (l + 1 >= n_levels))
            want = np.where(same, own, np.float32(np.nan)).astype(np.float32)
            finer = nbl < l
            if finer.any():
                tmp, cs = down[l]
                x, y, z = rx[finer], ry[finer], rz[finer]
                b = (x >> L) + (y >> L) * nbx + (z >> L) * nbx * nby
                want[finer] = cs[tmp.b2s[b], (x & (E - 1)) | ((y & (E - 1)) << L) | ((z & (E - 1)) << 2 * L)]
            coarser = ~same & ~finer
            c = lg.grids[min(l + 1, n_levels - 1)]
            for idx in zip(*np.nonzero(coarser)):
                want[idx] = orc.orc_sample_coarse(c.sx, c.sy, c.sz, c.L, lg.slabs[l + 1], c.b2s, int(rx[idx]),
                                                  int(ry[idx]), int(rz[idx]))
            checked[0] += int(same.sum())
            checked[1] += int(coarser.sum())
            checked[2] += int(finer.sum())
            bad = want.view(np.uint32) != tile.astype(np.float32).view(np.uint32)
            assert not bad.any(), (l, bid, np.argwhere(bad)[:3])
    assert min(checked.values()) > 100, checked


def _as_grid(g, slab):
    """A view of level grid g with `slab` as its payload (for orc.halo_fill)."""
    class V:
        pass
    v = V()
    v.sx, v.sy, v.sz, v.L, v.B = g.sx, g.sy, g.sz, g.L, 1 << g.L
    v.slab, v.b2s = slab, np.ascontiguousarray(g.b2s)
    return v


@pytest.mark.parametrize("stencil,dt", [(0, 0.1), (1, 0.02), (2, 0.005)])
def test_level_sweep_constant_field_stays_constant(pkg, orc, stencil, dt):
    lg = LevelGrid(pkg, orc, (96, 80, 64), 4, seed=3, const=0.625)
    active = np.arange(lg.levels.size)
    lg.sweep(orc, active, stencil, dt, 3)
    for l, a in enumerate(lg.slabs):
        assert np.all(a[2:] == np.float32(0.625)), l


@pytest.mark.parametrize("stencil", [0, 1, 2])
def test_level_sweep_of_fine_blocks_is_the_multires_sweep(pkg, orc, stencil):
    """With only level-0 blocks active the all-level sweep is the fine-only
    multires sweep (fill_multires halos), bit for bit; with coarse blocks
    active too, the coarse levels change and level 0 differs."""
    lg = LevelGrid(pkg, orc, (96, 80, 64), 3, seed=7 + stencil)
    fine = np.nonzero(lg.levels == 0)[0].astype(np.int64)
    want = lg.slabs[0].copy()
    f, c = lg.grids[0], lg.grids[1]
    orc.orc_sweep_multires(f.sx, f.sy, f.sz, 4, want, f.b2s, lg.slabs[1], c.b2s, lg.levels, fine, fine.size, stencil,
                           0.02, 2, 0, 1)
    before = [a.copy() for a in lg.slabs]
    lg.sweep(orc, fine, stencil, 0.02, 2)
    assert np.array_equal(lg.slabs[0].view(np.uint32), want.view(np.uint32))
    assert all(np.array_equal(a, b) for a, b in zip(lg.slabs[1:], before[1:]))
    lg.slabs = [b.copy() for b in before]
    lg.sweep(orc, np.arange(lg.levels.size), stencil, 0.02, 2)
    assert not np.array_equal(lg.slabs[1], before[1])
    assert not np.array_equal(lg.slabs[0], want)
