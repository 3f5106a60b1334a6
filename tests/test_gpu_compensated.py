"""GPU parity of the compensated iterate x + x_lo (two FP32 fields), through
the C ABI:

- msbg_poisson_residual_pair bit for bit against the oracle;
- refinement and multigrid solves with x_lo reach residuals an FP32 iterate
  cannot hold (||x|| / ||b|| ~ 1e4 here), and the residual they report equals
  the float64 residual of x + x_lo recomputed on the host;
- msbg_mac_subtract_gradient_pair bit for bit against the oracle;
- the FLIP step with a compensated pressure converges to 1e-6 where the plain
  FP32 pressure stalls at its floor, and its projection is the oracle's.
"""
import ctypes as C

import numpy as np
import pytest

from solve_helpers import band_problem, dense, listed_mask, neighbour_sum64
from test_gpu_sweep import assert_bitexact

pytestmark = pytest.mark.gpu


def big_solution_problem(pkg, s, seed):
    """A band problem whose solution is ~1e4 times its right-hand side:
    b = A x* in float64 (rounded to FP32) for a smooth x* of mean 3000."""
    g, active, x, _ = band_problem(pkg, *s, seed, frac=0.8, empty=0.0, full=0.0)
    z, y, xx = np.meshgrid(np.arange(g.sz), np.arange(g.sy), np.arange(g.sx), indexing="ij")
    X = 3000.0 + 50.0 * np.sin(0.21 * xx) * np.cos(0.17 * y) + 0.3 * z
    m = listed_mask(g, active)
    Xb = np.where(m, X, 0.0)  # Dirichlet 0 outside the listed cells
    # unlisted value blocks keep their stored values: make them 0 too
    xs = np.zeros((g.n_slots, 4096), np.float32)
    B = Xb * 6.0 - neighbour_sum64(Xb)
    b = np.zeros_like(xs)
    bid = (xx >> 4) + (y >> 4) * g.nx + (z >> 4) * g.nxy
    vid = (xx & 15) | ((y & 15) << 4) | ((z & 15) << 8)
    b[g.b2s[bid[m]], vid[m]] = B[m].astype(np.float32)
    return g, active, xs, b


def host_rel(g, active, hi, lo, b):
    X = dense(g, hi) + dense(g, lo)
    m = listed_mask(g, active)
    R = dense(g, b) - (6.0 * X - neighbour_sum64(X))
    return float(np.linalg.norm(R[m]) / np.linalg.norm(dense(g, b)[m]))


def test_residual_pair_bitexact(pkg, orc, hip):
    g, active, x, b = band_problem(pkg, 64, 48, 80, seed=3)
    rng = np.random.default_rng(3)
    x64 = 2000.0 + rng.normal(size=x.shape)
    hi = x64.astype(np.float32)
    lo = (x64 - hi).astype(np.float32)
    hi[:2], lo[:2] = x[:2], 0.0
    layout = pkg.DeviceGrid.from_host(g, lib=hip)
    band = pkg.PoissonBand(layout, active, lib=hip)
    r = band.field(0)
    r2, b2 = band.residual_pair(hip.to_device(hi), hip.to_device(lo), hip.to_device(b), r)
    want = np.zeros_like(hi)
    w2, wb2 = C.c_double(), C.c_double()
    orc.orc_poisson_residual_pair(g.sx, g.sy, g.sz, hi.ctypes.data, lo.ctypes.data, b.ctypes.data, g.b2s, active,
                                  active.size, want.ctypes.data, C.byref(w2), C.byref(wb2))
    m = np.zeros(g.n_slots, bool)
    m[g.b2s[active]] = True
    assert_bitexact(r.get().reshape(-1, 4096)[m], want[m], "pair residual")
    assert abs(r2 - w2.value) <= 1e-12 * w2.value and abs(b2 - wb2.value) <= 1e-12 * wb2.value


@pytest.mark.parametrize("storage,multigrid", [(1, True), (0, True), (1, False)])
def test_compensated_solve_goes_below_the_fp32_floor(pkg, hip, storage, multigrid):
    g, active, x0, b = big_solution_problem(pkg, (64, 48, 80), seed=5)
    layout = pkg.DeviceGrid.from_host(g, lib=hip)
    band = pkg.PoissonBand(layout, active, lib=hip)
    db = hip.to_device(b)
    kw = dict(multigrid=True, omega=1.0, sweeps_per_check=2, max_sweeps=400) if multigrid else \
        dict(refine=True, omega=1.9, sweeps_per_check=10, max_sweeps=20000)
    # plain FP32 iterate: stalls well above 1e-9
    x = band.field(0)
    st = band.solve(x, db, storage=storage, tol=1e-9, **kw)
    plain = host_rel(g, active, x.get().reshape(-1, 4096), np.zeros_like(x0), b)
    assert not st["converged"] and plain > 1e-8, (st, plain)
    # compensated pair
    x, lo = band.field(0), band.field(0)
    st = band.solve(x, db, storage=storage, tol=1e-9, x_lo=lo, **kw)
    hi_h, lo_h = x.get().reshape(-1, 4096), lo.get().reshape(-1, 4096)
    rel = host_rel(g, active, hi_h, lo_h, b)
    print("storage %d mg %d: plain floor %.3g, pair %.3g (reported %.3g) after %d sweeps"
          % (storage, multigrid, plain, rel, st["rel_residual"], st["sweeps"]))
    assert st["converged"] and st["rel_residual"] <= 1e-9
    assert rel <= 1.2e-9 and abs(rel - st["rel_residual"]) <= 0.05 * rel + 1e-12
    # the pair is normalised: |lo| <= ulp(hi) / 2
    m = np.zeros(g.n_slots, bool)
    m[g.b2s[active]] = True
    assert np.all(np.abs(lo_h[m]) <= 0.5 * np.spacing(np.abs(hi_h[m])))


def test_flip_step_compensated_pressure(pkg, orc, hip):
    """64^3 dam break: the compensated pressure converges to 1e-7 (the plain
    one stalls near 1.3e-6, its FP32 floor), and the projection of the GPU's
    u* with the GPU's (p, p_lo) is the oracle's, bit for bit."""
    from msbg_rs_amd import scenes
    from msbg_rs_amd.flip import FlipStep
    S = 64
    pos, vel, mass = scenes.dam_break(S, 200000)
    fs = FlipStep(S, S, S, capacity_blocks=(S // 16) ** 3, lib=hip)
    dp, dv, dm = hip.to_device(pos), hip.to_device(vel), hip.to_device(mass)
    _, _, st0 = fs.run(dp, dv, dm, storage=1, tol=1e-7, multigrid=True, omega=1.0, sweeps_per_check=2,
                       max_sweeps=400)
    assert not st0["solve"]["converged"]
    _, _, st = fs.run(dp, dv, dm, storage=1, tol=1e-7, multigrid=True, omega=1.0, sweeps_per_check=2,
                      max_sweeps=400, compensated=True)
    print("plain %.3g, compensated %.3g" % (st0["solve"]["rel_residual"], st["solve"]["rel_residual"]))
    assert st["solve"]["converged"] and st["solve"]["rel_residual"] <= 1e-7
    layout, band, p, b, active, band_blocks = fs.last
    L = layout.to_host()
    ut = [fs.mac_tilde.velocity[a].to_host() for a in range(3)]
    u, v, w = (np.ascontiguousarray(c.slab).copy() for c in ut)
    pl = np.ascontiguousarray(p.get().reshape(-1, 4096))
    plo = np.ascontiguousarray(fs.p_lo.get().reshape(-1, 4096))
    orc.orc_mac_subtract_gradient_pair(S, S, S, u, ut[0].b2s, v, ut[1].b2s, w, ut[2].b2s, pl, plo, L.b2s,
                                       active.astype(np.int64), active.size, -1.0)
    for a, want in enumerate((u, v, w)):
        assert_bitexact(fs.mac.velocity[a].to_host().slab[: ut[a].n_slots], want[: ut[a].n_slots], "axis %d" % a)
    n0, n1 = fs.divergence_norms()
    print("divergence %.4g -> %.4g" % (n0, n1))
    assert n1 <= 5e-6 * n0
