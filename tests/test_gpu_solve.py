"""GPU parity of the Poisson relaxation: storage conversions, Jacobi and
red-black sweeps, the residual with FP64 norms, and whole solves (FP32 and
FP16 / BF16 iterative refinement) against the oracle."""
import numpy as np
import pytest

from solve_helpers import (band_problem, dense, from_storage, listed_mask, oracle_cg64, oracle_jacobi, oracle_rbgs,
                           oracle_residual, oracle_solve, to_storage)
from test_gpu_sweep import assert_bitexact
from test_oracle_solve import convert, special_f32
from transfer_helpers import dense_from_slabs

pytestmark = pytest.mark.gpu


def setup(pkg, hip, g, active):
    layout = pkg.DeviceGrid.from_host(g, lib=hip)
    return layout, pkg.PoissonBand(layout, active, lib=hip)


def assert_same_bits(a, b, what):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    if a.dtype == np.uint16 or a.dtype == np.float16:
        ua, ub = a.view(np.uint16), b.view(np.uint16)
        if not np.array_equal(ua, ub):
            i = int(np.argmax(ua.ravel() != ub.ravel()))
            raise AssertionError("%s: %d differ; first at %d: %#x vs %#x" % (what, int((ua != ub).sum()), i,
                                                                         ua.ravel()[i], ub.ravel()[i]))
    else:
        assert_bitexact(a, b, what)


def test_storage_conversions_bitexact(pkg, orc, hip):
    rng = np.random.default_rng(0)
    f = special_f32(rng, 300000)
    f = f[~np.isnan(f)]
    d = hip.to_device(f)
    for st in (1, 2):
        out = hip.empty(f.size, np.uint16)
        hip.call("msbg_poisson_convert", 0, d.ptr, st, out.ptr, f.size)
        assert_same_bits(out.get(), convert(orc, f, 0, st), "f32 -> %d" % st)
        allv = np.arange(65536, dtype=np.uint32).astype(np.uint16)
        back = hip.empty(allv.size, np.float32)
        hip.call("msbg_poisson_convert", st, hip.to_device(allv).ptr, 0, back.ptr, allv.size)
        assert_bitexact(back.get(), convert(orc, allv, st, 0), "%d -> f32" % st)


# Small grids run the shallowest column kernels; (200, 160, 176) at 90% fill
# has >= 1024 listed blocks (FP32 columns of 2 cells), (250, 256, 244) >= 2048
# (FP32 columns of 4, 16-bit of 2), both with partial blocks at the far faces.
GRIDS = [((40, 33, 48), 0, 0.7), ((32, 32, 32), 1, 0.7), ((17, 50, 31), 2, 0.7), ((64, 48, 80), 3, 0.7),
         ((200, 160, 176), 6, 0.9), ((250, 256, 244), 7, 0.9)]


@pytest.mark.parametrize("s,seed,frac", GRIDS)
@pytest.mark.parametrize("st", [0, 1, 2])
def test_jacobi_bitexact(pkg, orc, hip, s, seed, frac, st):
    g, active, x, b = band_problem(pkg, *s, seed, frac=frac)
    _, pb = setup(pkg, hip, g, active)
    xs, bs = to_storage(x, st), to_storage(b, st)
    dx, db, dy = hip.to_device(xs), hip.to_device(bs), hip.to_device(xs)
    pb.jacobi(st, dx, dy, db, 2.0 / 3.0)
    assert_same_bits(dy.get(), oracle_jacobi(orc, g, active, xs, bs, 2.0 / 3.0, st), "jacobi")


@pytest.mark.parametrize("st", [0, 1, 2])
def test_jacobi_sweeps_graph_bitexact(pkg, orc, hip, st):
    """msbg_poisson_jacobi_sweeps: N graph-replayed sweeps x -> tmp -> x equal
    N oracle Jacobi applications; replayed twice back to back without a sync,
    then on a second pair of buffers (the graph is re-captured while the
    first may still run), and a different sweep count."""
    g, active, x, b = band_problem(pkg, 64, 48, 80, 3, frac=0.7)
    _, pb = setup(pkg, hip, g, active)
    xs, bs = to_storage(x, st), to_storage(b, st)
    db = hip.to_device(bs)
    want = xs
    for _ in range(20 + 20):
        want = oracle_jacobi(orc, g, active, want, bs, 2.0 / 3.0, st)
    dx, dt = hip.to_device(xs), hip.to_device(xs)
    pb.jacobi_sweeps(st, dx, dt, db, 2.0 / 3.0, 20)
    pb.jacobi_sweeps(st, dx, dt, db, 2.0 / 3.0, 20)
    ex, et = hip.to_device(xs), hip.to_device(xs)  # new buffers: re-capture
    pb.jacobi_sweeps(st, ex, et, db, 2.0 / 3.0, 20)
    pb.jacobi_sweeps(st, ex, et, db, 2.0 / 3.0, 12)
    hip.synchronize()
    assert_same_bits(dx.get(), want, "40 graph sweeps")
    want2 = xs
    for _ in range(32):
        want2 = oracle_jacobi(orc, g, active, want2, bs, 2.0 / 3.0, st)
    assert_same_bits(ex.get(), want2, "20 + 12 graph sweeps on new buffers")


@pytest.mark.parametrize("s,seed,frac", GRIDS)
@pytest.mark.parametrize("st", [0, 1, 2])
def test_rbgs_bitexact(pkg, orc, hip, s, seed, frac, st):
    g, active, x, b = band_problem(pkg, *s, seed + 10, frac=frac)
    _, pb = setup(pkg, hip, g, active)
    xs, bs = to_storage(x, st), to_storage(b, st)
    dx, db = hip.to_device(xs), hip.to_device(bs)
    want = xs
    for it in range(3):
        for color in (0, 1):
            pb.rbgs(st, dx, db, color, 1.9)
            want = oracle_rbgs(orc, g, active, want, bs, color, 1.9, st)
    assert_same_bits(dx.get(), want, "rbgs")


@pytest.mark.parametrize("s,frac", [((64, 48, 80), 0.7), ((200, 160, 176), 0.9), ((250, 256, 244), 0.9)])
@pytest.mark.parametrize("st", [0, 1, 2])
def test_residual_bitexact_and_norms(pkg, orc, hip, st, s, frac):
    g, active, x, b = band_problem(pkg, *s, 4, frac=frac)
    _, pb = setup(pkg, hip, g, active)
    xs, bs = to_storage(x, st), to_storage(b, st)
    r = pb.field(0)
    r2, b2 = pb.residual(st, hip.to_device(xs), hip.to_device(bs), r)
    ro, r2o, b2o = oracle_residual(orc, g, active, xs, bs, st)
    m = listed_mask(g, active)
    assert_bitexact(dense(g, r.get())[m].astype(np.float32), dense(g, ro)[m].astype(np.float32), "residual")
    # Both norms sum the same non-negative FP64 terms in different orders;
    # each sum is within (n - 1) u of the exact one (u = 2^-53).
    rel = max(1e-12, 2.0 * int(m.sum()) * 2.0 ** -53)
    assert r2 == pytest.approx(r2o, rel=rel) and b2 == pytest.approx(b2o, rel=rel)


@pytest.mark.parametrize("s,frac,sweeps", [((48, 40, 33), 0.7, 40), ((250, 256, 244), 0.9, 10)])
@pytest.mark.parametrize("storage,refine", [(0, 0), (0, 1), (1, 1), (2, 1)])
def test_solve_fixed_sweeps_bitexact(pkg, orc, hip, storage, refine, s, frac, sweeps):
    """tol = 0: exactly max_sweeps sweeps on both sides; x must agree bit for bit."""
    g, active, x, b = band_problem(pkg, *s, 5, frac=frac, empty=0.0, full=0.0)
    _, pb = setup(pkg, hip, g, active)
    x0 = np.zeros_like(x)
    dx, db = hip.to_device(x0), hip.to_device(b)
    st = pb.solve(dx, db, storage=storage, refine=bool(refine), omega=1.9, sweeps_per_check=10, max_sweeps=sweeps,
                  tol=0.0)
    xo, sto = oracle_solve(orc, g, active, x0, b, storage, refine, omega=1.9, per=10, max_sweeps=sweeps, tol=0.0)
    assert st["sweeps"] == sto["sweeps"] == sweeps and st["checks"] == sto["checks"] == sweeps // 10 + 1
    assert_bitexact(dx.get(), xo, "solve x")
    # FP64 norms summed in different orders (see test_residual_bitexact_and_norms)
    rel = max(1e-12, 2.0 * int(listed_mask(g, active).sum()) * 2.0 ** -53)
    assert st["rel_residual"] == pytest.approx(sto["rel_residual"], rel=rel)


@pytest.mark.parametrize("storage,refine", [(0, 0), (1, 1), (2, 1)])
def test_shell_solve_to_1e6_vs_fp64(pkg, orc, hip, storage, refine):
    """The 128^3 shell band (absent blocks: zero Dirichlet), LCG right-hand
    side: the GPU solve reaches ||r|| / ||b|| <= 1e-6 and its error against
    the FP64 CG solution is small."""
    from msbg_rs_amd import scenes
    g, bids, _ = scenes.shell_scene()
    _, pb = setup(pkg, hip, g, bids)
    b = scenes.poisson_rhs(g, bids)
    dx, db = pb.field(0), hip.to_device(b)
    st = pb.solve(dx, db, storage=storage, refine=bool(refine), omega=1.9, sweeps_per_check=10, max_sweeps=3000,
                  tol=1e-6)
    assert st["converged"] and st["rel_residual"] <= 1e-6, st
    xg = dx.get().reshape(-1, 4096)
    _, r2, b2 = oracle_residual(orc, g, bids, np.ascontiguousarray(xg), b)
    assert np.sqrt(r2 / b2) <= 1e-6
    xd, _ = oracle_cg64(orc, g, bids, b.astype(np.float64), tol=1e-12)
    m = listed_mask(g, bids)
    s = (g.sx, g.sy, g.sz)
    ref = dense_from_slabs(g.b2s, xd, s)[m]
    err = np.abs(dense_from_slabs(g.b2s, xg, s)[m] - ref).max() / np.abs(ref).max()
    assert err < 1e-5, err
