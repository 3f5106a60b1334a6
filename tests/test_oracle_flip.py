"""Pins of the oracle's pressure projection of MAC face velocities
(orc_mac_subtract_gradient; no reference counterpart, the reference stops at
the divergence, fluid/mac.rs:134-171):

- a linear pressure p = x + 2y + 3z on whole blocks: every face moves by
  exactly -scale * (1, 2, 3), the faces on the domain's far walls by 0;
- the discrete identity the projection exists for: with scale = -1 the new
  divergence is div - A p (A the band operator of orc_poisson_residual: the
  residual of A p = div) on every cell whose three lower neighbours are
  stored, for any p.
"""
import numpy as np

from test_oracle_transfer import mac_channels, oracle_divergence


def project(orc, chans, pgrid, pslab, active, scale):
    s = (chans[0].sx, chans[0].sy, chans[0].sz)
    u, v, w = (np.ascontiguousarray(c.slab).copy() for c in chans)
    orc.orc_mac_subtract_gradient(*s, u, chans[0].b2s, v, chans[1].b2s, w, chans[2].b2s,
                                  np.ascontiguousarray(pslab, np.float32), pgrid.b2s,
                                  np.ascontiguousarray(active, np.int64), len(active), scale)
    return u, v, w


def test_linear_pressure_moves_faces_exactly(pkg, orc):
    s = (48, 32, 32)
    f = lambda a, X, Y, Z: np.sin(0.3 * X + a) + 0.25 * Y - 0.5 * Z
    chans = mac_channels(pkg, s, f)
    pg = pkg.SparseGrid("p", *s, 0.0, 0.0)
    act = np.arange(pg.n_blocks)
    pg.ensure_blocks_parallel(act)
    pg.fill_from_function(act, lambda X, Y, Z: (X + 2.0 * Y + 3.0 * Z).astype(np.float32))
    for scale in (1.0, -0.5):
        u, v, w = project(orc, chans, pg, pg.slab, act, scale)
        for a, (new, c) in enumerate(zip((u, v, w), chans)):
            old = c.to_dense()
            g = pkg.SparseGrid("n", *s, 0.0, 0.0)
            g.b2s, g._slab, g.n_slots = c.b2s, new.reshape(-1, 4096), c.n_slots
            d = g.to_dense().astype(np.float64) - old.astype(np.float64)
            want = np.full(d.shape, -scale * (a + 1.0))
            edge = [slice(None)] * 3
            edge[2 - a] = slice(s[a] - 1, s[a])  # dense arrays are (z, y, x)
            want[tuple(edge)] = 0.0
            # |face| < 64 so the float32 difference of the update is exact to 1e-5
            np.testing.assert_allclose(d, want, atol=1e-5)


def test_projection_identity_with_band_residual(pkg, orc):
    from solve_helpers import oracle_residual
    s = (40, 33, 48)
    f = lambda a, X, Y, Z: np.sin(0.21 * X + 0.4 * a) * np.cos(0.13 * Y) + 0.03 * Z * (a - 1)
    chans = mac_channels(pkg, s, f)
    act = np.arange(chans[0].n_blocks)
    div = oracle_divergence(pkg, orc, chans, act)
    rng = np.random.default_rng(0)
    p = rng.normal(size=div.slab.shape).astype(np.float32)
    p[:2] = 0.0  # absent / full dummies read 0
    u, v, w = project(orc, chans, div, p, act, -1.0)
    new = []
    for c, x in zip(chans, (u, v, w)):
        g = pkg.SparseGrid("n", *s, 0.0, 0.0)
        g.b2s, g._slab, g.n_slots = c.b2s, x.reshape(-1, 4096), c.n_slots
        new.append(g)
    div_new = oracle_divergence(pkg, orc, new, act).to_dense().astype(np.float64)
    r, _, _ = oracle_residual(orc, div, act.astype(np.int64), p, np.ascontiguousarray(div.slab, np.float32), 0)
    g = pkg.SparseGrid("r", *s, 0.0, 0.0)
    g.b2s, g._slab, g.n_slots = div.b2s, np.asarray(r, np.float32).reshape(-1, 4096), div.n_slots
    rd = g.to_dense().astype(np.float64)
    # every block is stored, so the identity holds on every cell up to rounding
    scale = np.abs(div.to_dense()).max() + 12 * np.abs(p).max()
    assert np.abs(div_new - rd).max() <= 2e-6 * scale
    assert np.abs(rd).max() > 1.0  # p is no solution: the identity is not trivially 0 = 0
