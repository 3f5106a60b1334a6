"""Pins of the oracle's density min-splat (particles/splat.rs:92-319): the
reference's splat_01..06 known answers (splat.rs:371-470) and a brute-force
per-particle formulation with the same region rules."""
import numpy as np
import pytest

R_SCAN = 4.0  # r_particle 2 + nb_dist 2: the tests' and the demo's r_scan (particles/mod.rs:51-52)
FULL = 65535


def u16p():
    return np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")


@pytest.fixture(scope="module")
def splat(orc):
    import ctypes as C
    fn = orc.lib.orc_density_splat
    fn.restype = None
    fn.argtypes = [C.c_int64] * 3 + [np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS"), u16p(),
                                     np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS"), C.c_int64, C.c_float,
                                     C.c_int]
    return fn


def density_grid(pkg, orc, s, pos, r_scan=R_SCAN):
    """b2s with value blocks on the particles' active blocks (active.rs:15-47),
    and a u16 slab with those blocks filled full (the caller's fill_active)."""
    g = pkg.SparseGrid("density", *s, 0.0, 1.0)
    act = orc.active_blocks(*s, pos, r_scan)
    g.ensure_blocks_parallel(act)
    slab = np.zeros((g.n_slots, 4096), np.uint16)
    slab[1] = FULL
    slab[g.b2s[act]] = FULL
    return g, slab


def run(splat, g, slab, pos, r_scan=R_SCAN, ref=0):
    out = slab.copy()
    p = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    splat(g.sx, g.sy, g.sz, g.b2s, out.reshape(-1), p.reshape(-1), p.shape[0], r_scan, ref)
    return out


def voxel(g, slab, x, y, z):
    return int(slab[g.b2s[g.get_block_id(x, y, z)], (x & 15) | ((y & 15) << 4) | ((z & 15) << 8)])



def test_splat_01_06_single_particle_shape(pkg, orc, splat):
    pos = np.array([[8.0, 8.0, 8.0]], np.float32)
    g, slab = density_grid(pkg, orc, (32, 32, 32), pos)
    out = run(splat, g, slab, pos)
    assert voxel(g, out, 8, 8, 8) == 3072  # distSq 0.75
    assert voxel(g, out, 5, 8, 8) == round((6.75 / 16.0) * 65535.0)
    assert voxel(g, out, 4, 8, 8) == round((12.75 / 16.0) * 65535.0)
    assert voxel(g, out, 1, 8, 8) == FULL and voxel(g, out, 0, 8, 8) == FULL


def test_splat_02_boundary_crossing_particle(pkg, orc, splat):
    pos = np.array([[15.5, 8.0, 8.0]], np.float32)
    g, slab = density_grid(pkg, orc, (32, 32, 32), pos)
    assert g.b2s[1] >= 2
    out = run(splat, g, slab, pos)
    assert voxel(g, out, 16, 8, 8) == round((1.5 / 16.0) * 65535.0)


def test_splat_03_overlap_min_is_order_independent(pkg, orc, splat):
    pos = np.array([[10.0, 8.0, 8.0], [12.0, 8.0, 8.0]], np.float32)
    g, slab = density_grid(pkg, orc, (32, 32, 32), pos)
    a = run(splat, g, slab, pos)
    b = run(splat, g, slab, pos[::-1].copy())
    assert np.array_equal(a, b)


def test_splat_04_05_corner_clip_and_radius_cull(pkg, orc, splat):
    pos = np.array([[0.1, 0.1, 0.1]], np.float32)
    g, slab = density_grid(pkg, orc, (32, 32, 32), pos)
    assert voxel(g, run(splat, g, slab, pos), 0, 0, 0) != FULL
    pos = np.array([[16.0, 16.0, 16.0]], np.float32)
    g, slab = density_grid(pkg, orc, (32, 32, 32), pos)
    out = run(splat, g, slab, pos)
    assert voxel(g, out, 21, 16, 16) == FULL and voxel(g, out, 20, 16, 16) == FULL


def brute_force(g, slab, pos, r_scan):
    """Per particle, per footprint voxel: the staged value, committed when the
    voxel's region (relative to the particle's block) lies inside the domain
    and targets a value block."""
    out = slab.copy()
    m = int(np.ceil(r_scan))
    f = np.float32
    dmax, inv = f(r_scan) * f(r_scan), f(1.0) / (f(r_scan) * f(r_scan))
    s = (g.sx, g.sy, g.sz)
    for p in np.asarray(pos, np.float32):
        b0 = [int(max(float(c), 0.0)) >> 4 << 4 for c in p]

        def axis(k):
            lo = max(int(np.ceil(f(f(p[k]) - f(r_scan)) - f(0.5))), b0[k] - m)
            hi = min(int(np.floor(f(f(p[k]) + f(r_scan)) - f(0.5))), b0[k] + 15 + m)
            return np.arange(lo, hi + 1)

        X, Y, Z = np.meshgrid(axis(0), axis(1), axis(2), indexing="ij")
        dz = Z.astype(f) - f(p[2] - f(0.5))
        dy = Y.astype(f) - f(p[1] - f(0.5))
        dx = X.astype(f) - f(p[0] - f(0.5))
        d2 = (dx * dx) + ((dz * dz) + (dy * dy))
        keep = d2 <= dmax
        qv = np.floor((d2 * inv * f(65535.0)).astype(np.float64) + 0.5).astype(np.int64)
        for x, y, z, v in zip(X[keep], Y[keep], Z[keep], qv[keep]):
            ok = True
            for c, k in ((x, 0), (y, 1), (z, 2)):
                lo = b0[k] - m if c < b0[k] else (b0[k] if c < b0[k] + 16 else b0[k] + 16)
                width = m if (c < b0[k] or c >= b0[k] + 16) else 16
                ok &= lo >= 0 and lo + width <= s[k]
            if not ok:
                continue
            slot = g.b2s[g.get_block_id(x, y, z)]
            if slot < 2:
                continue
            i = (x & 15) | ((y & 15) << 4) | ((z & 15) << 8)
            out[slot, i] = min(out[slot, i], v)
    return out


@pytest.mark.parametrize("r_scan,s", [(4.0, (48, 40, 33)), (2.5, (40, 33, 48)), (3.2, (32, 32, 32))])
def test_staged_splat_equals_brute_force(pkg, orc, splat, r_scan, s):
    rng = np.random.default_rng(int(r_scan * 10))
    pos = rng.uniform(0.0, np.array(s) - 1e-3, size=(600, 3)).astype(np.float32)
    pos[:40] = np.round(pos[:40] / 16) * 16 + rng.uniform(-0.6, 0.6, size=(40, 3)).astype(np.float32)  # block faces
    pos = np.clip(pos, 0.0, np.array(s, np.float32) - 1e-3).astype(np.float32)
    g, slab = density_grid(pkg, orc, s, pos, r_scan)
    got = run(splat, g, slab, pos, r_scan)
    assert np.array_equal(got, brute_force(g, slab, pos, r_scan))


def test_reference_margin_commit_width(pkg, orc, splat):
    """The reference commits margin rows 4 lanes wide (splat.rs:306-309): equal
    to the exact width for r_scan in (3, 4]; with m = 2 it also writes the two
    interior staging columns into the next row of the -x neighbour."""
    pos = np.array([[16.3, 8.2, 8.4]], np.float32)
    g, slab = density_grid(pkg, orc, (48, 32, 32), pos, 4.0)
    assert np.array_equal(run(splat, g, slab, pos, 4.0, 1), run(splat, g, slab, pos, 4.0, 0))
    g, slab = density_grid(pkg, orc, (48, 32, 32), pos, 2.0)
    exact, ref = run(splat, g, slab, pos, 2.0, 0), run(splat, g, slab, pos, 2.0, 1)
    diff = np.argwhere(exact != ref)
    assert diff.size and np.all(g.b2s[0] == diff[:, 0])  # only the -x neighbour (block 0) differs
    vid = diff[:, 1]
    assert np.all(vid & 15 < 2)  # x = 0, 1 of rows past the margin rows
