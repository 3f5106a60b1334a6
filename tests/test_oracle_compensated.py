"""Pins of the oracle's compensated-iterate functions (no reference
counterpart; the FP32 pair x + x_lo of the refinement solve):

- orc_poisson_residual_pair equals the residual b - A (x + x_lo) evaluated
  in float64 (exact up to 2^-53 of the largest term) within half an FP32 ulp
  of the result plus 2^-40 of the largest term; a plain FP32 residual of x
  misses that bound by orders of magnitude, and so does the pair residual with
  its fma split of 6 x dropped (planted below);
- with x_lo = 0 its bits equal the plain residual's wherever the hi sums are
  exact (small integers);
- orc_mac_subtract_gradient_pair with p_lo = 0 is bit-identical to
  orc_mac_subtract_gradient, and with a float64 linear pressure split into
  hi + lo it moves every face by the float64 gradient to within an ulp of the
  face.
"""
import ctypes as C

import numpy as np

from solve_helpers import band_problem, dense, listed_mask, neighbour_sum64
from test_oracle_flip import project
from test_oracle_transfer import mac_channels


def residual_pair(orc, g, active, x, lo, b):
    r = np.zeros(x.shape, np.float32)
    r2, b2 = C.c_double(), C.c_double()
    orc.orc_poisson_residual_pair(g.sx, g.sy, g.sz, x.ctypes.data, lo.ctypes.data, b.ctypes.data, g.b2s, active,
                                  active.size, r.ctypes.data, C.byref(r2), C.byref(b2))
    return r, r2.value, b2.value


def split_pair(v64):
    hi = v64.astype(np.float32)
    lo = (v64 - hi.astype(np.float64)).astype(np.float32)
    return hi, lo


def test_pair_residual_matches_float64(pkg, orc):
    g, active, x, b = band_problem(pkg, 40, 33, 48, seed=4)
    rng = np.random.default_rng(4)
    # a large iterate (||x|| >> ||b||, where one FP32 field's floor shows)
    x64 = 1000.0 + rng.normal(size=x.shape)
    x64[0], x64[1] = 0.25, 0.75  # the dummy slots as stored
    hi, lo = split_pair(x64)
    lo[:2] = 0.0
    r, r2, b2 = residual_pair(orc, g, active, hi, lo, b)
    X = dense(g, hi) + dense(g, lo)
    R64 = dense(g, b) - (6.0 * X - neighbour_sum64(X))
    m = listed_mask(g, active)
    got = dense(g, r)[m]
    want = R64[m]
    bound = 0.5 * np.abs(want) * 2.0 ** -23 + 2.0 ** -40 * 7000.0
    assert np.all(np.abs(got - want) <= bound), np.abs(got - want).max()
    assert abs(r2 - float(np.sum(got * got))) <= 1e-9 * r2
    # the plain FP32 residual of hi alone is far outside that bound
    from solve_helpers import oracle_residual
    r32, _, _ = oracle_residual(orc, g, active, hi, b)
    assert (np.abs(dense(g, r32)[m] - want) / bound).max() > 1e3


def test_pair_residual_with_zero_low_part_is_the_plain_one_on_integers(pkg, orc):
    from solve_helpers import oracle_residual
    g, active, x, b = band_problem(pkg, 32, 32, 32, seed=1)
    rng = np.random.default_rng(1)
    xi = rng.integers(-50, 50, size=x.shape).astype(np.float32)
    bi = rng.integers(-50, 50, size=b.shape).astype(np.float32)
    xi[0], xi[1] = 0.0, 1.0
    r, _, _ = residual_pair(orc, g, active, xi, np.zeros_like(xi), bi)
    r0, _, _ = oracle_residual(orc, g, active, xi, bi)
    m = listed_mask(g, active)
    assert np.array_equal(dense(g, r)[m], dense(g, r0)[m])


def test_planted_missing_fma_split_fails_the_pin(pkg, orc):
    """The same evaluation in float32 without the exact split of 6 x (what a
    dropped error term gives) misses the float64 residual by far more than
    the pinned bound: the pin above would catch it."""
    g, active, x, b = band_problem(pkg, 32, 32, 32, seed=2)
    rng = np.random.default_rng(2)
    x64 = 3000.0 + rng.normal(size=x.shape)
    hi, lo = split_pair(x64)
    X = dense(g, hi).astype(np.float64) + dense(g, lo)
    R64 = dense(g, b) - (6.0 * X - neighbour_sum64(X))
    Xh = dense(g, hi).astype(np.float32)
    planted = (dense(g, b).astype(np.float32) - np.float32(6.0) * Xh) + neighbour_sum64(Xh).astype(np.float32)
    m = listed_mask(g, active)
    bound = 0.5 * np.abs(R64[m]) * 2.0 ** -23 + 2.0 ** -40 * 20000.0
    assert (np.abs(planted[m] - R64[m]) / bound).max() > 100


def test_pair_projection(pkg, orc):
    s = (48, 32, 32)
    f = lambda a, X, Y, Z: np.sin(0.3 * X + a) + 0.25 * Y - 0.5 * Z
    chans = mac_channels(pkg, s, f)
    pg = pkg.SparseGrid("p", *s, 0.0, 0.0)
    act = np.arange(pg.n_blocks)
    pg.ensure_blocks_parallel(act)
    rng = np.random.default_rng(0)
    pslab = rng.normal(size=(pg.n_slots, 4096)).astype(np.float32)
    pslab[:2] = 0.0
    want = project(orc, chans, pg, pslab, act, -1.0)

    def project_pair(p_hi, p_lo, scale):
        u, v, w = (np.ascontiguousarray(c.slab).copy() for c in chans)
        orc.orc_mac_subtract_gradient_pair(*s, u, chans[0].b2s, v, chans[1].b2s, w, chans[2].b2s,
                                           np.ascontiguousarray(p_hi), np.ascontiguousarray(p_lo), pg.b2s,
                                           act.astype(np.int64), act.size, scale)
        return u, v, w

    got = project_pair(pslab, np.zeros_like(pslab), -1.0)
    for a in range(3):
        assert np.array_equal(got[a].view(np.uint32), want[a].view(np.uint32))
    # a linear pressure 1000 + x/3 + 2y/7 + 3z/11 in float64, split hi + lo
    z, y, x = np.meshgrid(np.arange(32), np.arange(32), np.arange(48), indexing="ij")
    P = 1000.0 + x / 3.0 + 2.0 * y / 7.0 + 3.0 * z / 11.0
    hi = np.zeros((pg.n_slots, 4096), np.float32)
    lo = np.zeros_like(hi)
    bid = (x >> 4) + (y >> 4) * pg.nx + (z >> 4) * pg.nxy
    vid = (x & 15) | ((y & 15) << 4) | ((z & 15) << 8)
    h64, l64 = split_pair(P)
    hi[pg.b2s[bid], vid] = h64
    lo[pg.b2s[bid], vid] = l64
    out = project_pair(hi, lo, 1.0)
    grads = (1.0 / 3.0, 2.0 / 7.0, 3.0 / 11.0)
    for a in range(3):
        base = np.ascontiguousarray(chans[a].slab)
        moved = (base.astype(np.float64) - out[a].astype(np.float64))[2:]
        # interior faces (not on the far wall of the axis)
        coord = (x, y, z)[a]
        inner = coord < (s[a] - 1)
        sel = np.zeros_like(base, bool)
        sel[chans[a].b2s[bid[inner]], vid[inner]] = True
        d = np.abs(moved[sel[2:]] - grads[a])
        ulp = np.spacing(np.abs(base[sel]).astype(np.float32)).astype(np.float64) + 2.0 ** -24
        assert np.all(d <= 1.01 * ulp), d.max()
