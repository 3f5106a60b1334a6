"""Pins of the oracle's Density8 (u8) sweep (Sweeper<Density8>, solver.rs:
154-240 with StoreBack<Density8>, math/stencil.rs:176-192).

The reference's stochastic-rounding stream depends on its SIMD width and
thread schedule, so no reference vector exists; the stored codes are pinned
by what the quantiser fixes: each code is floor(vf) or floor(vf) + 1 of
vf = sqrt(v) * 255 (one more exactly when the draw is below the fraction),
the mean over draws is vf (unbiased), the draws are uniform and uncorrelated,
0 and 1 are exact, and the linear dequantisation is code * (1/255)
(math/sample.rs:64-69)."""
import numpy as np
import pytest

from test_gpu_sweep import random_grid

f32 = np.float32


def vf32(v):
    return np.sqrt(np.asarray(v, f32)) * f32(255.0)


def codes(orc, v, u):
    return np.array([orc.orc_density8_code(float(a), float(b)) for a, b in zip(v, u)], np.int64)


def draws(orc, seed, sweep, bids, vids):
    return np.array([orc.orc_density8_uniform(seed, sweep, int(b), int(c)) for b, c in zip(bids, vids)], f32)


def test_code_is_floor_or_next_by_the_draw(orc):
    rng = np.random.default_rng(0)
    v = rng.uniform(0, 1, 20000).astype(f32)
    v[:4] = [0.0, 1.0, 0.25, np.float32(180 / 255.0) ** 2]
    u = draws(orc, 7, 0, rng.integers(0, 1 << 20, v.size), rng.integers(0, 4096, v.size))
    q = codes(orc, v, u)
    x = vf32(v)
    fl = np.floor(x)
    assert np.all((q == fl) | (q == fl + 1))
    assert np.array_equal(q == fl + 1, u < (x - fl))
    assert q[0] == 0 and q[1] == 255 and q[2] in (127, 128)


def test_codes_are_unbiased_over_draws(orc):
    n = 40000
    for v in (0.01, 0.3, 0.5, 0.77, 0.999):
        u = draws(orc, 3, 1, np.arange(n) // 4096, np.arange(n) % 4096)
        q = codes(orc, [v] * n, u)
        x = float(vf32(v))
        fr = x - np.floor(x)
        sigma = np.sqrt(max(fr * (1 - fr), 1e-12) / n)
        assert abs(q.mean() - x) < 5 * sigma + 1e-9, (v, q.mean(), x)


def test_draws_are_uniform_and_uncorrelated(orc):
    n = 4096 * 16
    bids, vids = np.arange(n) // 4096, np.arange(n) % 4096
    a = draws(orc, 11, 0, bids, vids).astype(np.float64)
    b = draws(orc, 11, 1, bids, vids).astype(np.float64)  # next sweep
    c = draws(orc, 12, 0, bids, vids).astype(np.float64)  # other seed
    assert a.min() >= 0.0 and a.max() < 1.0
    assert abs(a.mean() - 0.5) < 0.005 and abs(a.var() - 1 / 12) < 0.002
    hist = np.histogram(a, bins=16, range=(0, 1))[0]
    assert hist.min() > 0.9 * n / 16 and hist.max() < 1.1 * n / 16
    for x, y in ((a[:-1], a[1:]), (a, b), (a, c)):
        assert abs(np.corrcoef(x, y)[0, 1]) < 0.02
    assert len(np.unique(a)) > 0.99 * n


def test_linear_dequantisation_of_all_codes(pkg, orc):
    """The halo value the kernels read is code * (1/255): in a field of
    constant code c the colour-0 blocks (which see no updated neighbour) are
    smoothed by the Laplacian to exactly c / 255 and store SR(sqrt(c / 255)
    * 255); the other blocks are untouched."""
    g = random_grid(pkg, 64, 64, 64, seed=1, frac=1.0, n_full=0, lo=0.0, hi=1.0)
    b = np.arange(g.n_blocks)
    active = b[((b % 4) % 2 == 0) & (((b // 4) % 4) % 2 == 0) & ((b // 16) % 2 == 0)].astype(np.int64)
    for c in (0, 1, 64, 128, 200, 255):
        q = np.full((g.n_slots, 4096), c, np.uint8)
        orc.orc_sweep_density8(g.sx, g.sy, g.sz, q, g.n_slots, g.b2s, active, active.size, 0, 0.1, 1, 1, 5, 0)
        x = float(vf32(f32(c) * f32(1.0 / 255.0)))
        untouched = np.ones(g.n_slots, bool)
        untouched[g.b2s[active]] = False
        assert np.all(q[untouched] == c)
        got = q[g.b2s[active]]
        fl = np.floor(x)
        assert np.all((got == fl) | (got == min(fl + 1, 255))), c
        if c in (0, 255):
            assert np.all(got == c)
        else:
            fr = x - fl
            assert abs(got.mean() - x) < 5 * np.sqrt(fr * (1 - fr) / got.size) + 1e-9, c


@pytest.mark.parametrize("stencil,dt", [(0, 0.1), (1, 0.02), (2, 0.005)])
def test_sweep_codes_bracket_the_unquantised_update(pkg, orc, stencil, dt):
    """One sweep of a single colour-0 block (every other block untouched):
    each stored code brackets sqrt(v) * 255 of the oracle's FP32 kernel output
    on the dequantised tile."""
    g = random_grid(pkg, 48, 48, 48, seed=4, frac=1.0, n_full=0, lo=0.0, hi=1.0)
    rng = np.random.default_rng(9)
    q = rng.integers(0, 256, size=(g.n_slots, 4096)).astype(np.uint8)
    q[0], q[1] = 0, 255
    bid = np.int64(0)
    fl32 = (q.astype(f32) * f32(1.0 / 255.0)).astype(f32)
    h = g.copy()
    h._slab[: h.n_slots] = fl32
    halo, full = (2, True) if stencil == 2 else (1, stencil == 1)
    want_v = orc.kernel(stencil, orc.halo_fill(h, int(bid), halo, full), dt, True)
    out = q.copy()
    orc.orc_sweep_density8(g.sx, g.sy, g.sz, out, g.n_slots, g.b2s, np.array([bid]), 1, stencil, dt, 1, 1, 3, 0)
    x = vf32(want_v).ravel()
    got = out[g.b2s[bid]].astype(np.int64)
    fl = np.floor(x)
    assert np.all((got == fl) | (got == np.minimum(fl + 1, 255)))
    others = np.ones(g.n_slots, bool)
    others[g.b2s[bid]] = False
    assert np.array_equal(out[others], q[others])
