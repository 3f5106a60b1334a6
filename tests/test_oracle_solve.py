"""Pins of the oracle's Poisson relaxation (no reference implementation
exists: docs/roadmap.md:610-702).  Each check is independent of the oracle's
own formulas: numpy / torch / scipy library routines, an FP64 dense statement
of the operator, closed-form eigenmodes and fixed points, and a serial
Gauss-Seidel in exact float32 arithmetic visiting cells in random order."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from transfer_helpers import dense_from_slabs

from solve_helpers import (addf, band_problem, dense, fmaf, from_storage, listed_mask, neighbour_sum64,
                           oracle_cg64, oracle_jacobi, oracle_rbgs, oracle_residual, oracle_solve, to_storage)


def convert(orc, src, st_src, st_dst):
    src = np.ascontiguousarray(src)
    out = np.zeros(src.size, {0: np.float32, 1: np.float16, 2: np.uint16}[st_dst])
    orc.orc_convert(st_src, src.ctypes.data, st_dst, out.ctypes.data, src.size)
    return out


def special_f32(rng, n):
    """Random float32 bit patterns biased to the f16 / bf16 edge cases."""
    u = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    e = rng.integers(100, 146, n).astype(np.uint32)  # exponents around the f16 subnormal / overflow range
    u[: n // 2] = (u[: n // 2] & 0x807FFFFF) | (e[: n // 2] << 23)
    u[n // 2: n // 2 + 1000] &= 0xFFFFE000  # exact f16 mantissas
    u[n // 2 + 1000: n // 2 + 2000] = (u[n // 2 + 1000: n // 2 + 2000] & 0xFFFFE000) | 0x1000  # ties
    u[n // 2 + 2000: n // 2 + 3000] = (u[n // 2 + 2000: n // 2 + 3000] & 0xFFFF0000) | 0x8000  # bf16 ties
    return u.view(np.float32)


def test_f16_conversion_is_ieee_rne(orc):
    rng = np.random.default_rng(0)
    f = special_f32(rng, 400000)
    f = f[~np.isnan(f)]
    got = convert(orc, f, 0, 1)
    with np.errstate(over="ignore"):
        want = f.astype(np.float16)  # numpy: IEEE binary16, round to nearest even
    assert np.array_equal(got.view(np.uint16), want.view(np.uint16))
    # every binary16 value converts to float32 exactly and back unchanged
    allh = np.arange(65536, dtype=np.uint32).astype(np.uint16).view(np.float16)
    allh = allh[~np.isnan(allh)]
    f32 = convert(orc, allh, 1, 0)
    assert np.array_equal(f32, allh.astype(np.float32))
    assert np.array_equal(convert(orc, f32, 0, 1).view(np.uint16), allh.view(np.uint16))


def test_bf16_conversion_is_rne(orc):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    f = special_f32(rng, 400000)
    f = f[~np.isnan(f)]
    got = convert(orc, f, 0, 2)
    want = torch.from_numpy(f.copy()).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(got, want)
    # closed-form ties: 1 + 2^-8 is halfway between 1 and 1 + 2^-7 -> even (1);
    # 1 + 3 * 2^-8 -> 1 + 2^-6 (even); just above the tie rounds up.
    t = np.array([1 + 2 ** -8, 1 + 3 * 2 ** -8, np.nextafter(np.float32(1 + 2 ** -8), np.float32(2))], np.float32)
    assert list(convert(orc, t, 0, 2)) == [0x3F80, 0x3F82, 0x3F81]
    allb = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    f32 = convert(orc, allb, 2, 0)
    ok = ~np.isnan(f32)
    assert np.array_equal(f32.view(np.uint32)[ok], allb.astype(np.uint32)[ok] << 16)
    assert np.array_equal(convert(orc, f32[ok], 0, 2), allb[ok])


@pytest.mark.parametrize("s,seed", [((40, 33, 48), 0), ((32, 32, 32), 1), ((17, 50, 31), 2)])
@pytest.mark.parametrize("st", [0, 1, 2])
def test_jacobi_matches_dense_fp64(pkg, orc, s, seed, st):
    g, active, x, b = band_problem(pkg, *s, seed)
    xs, bs = to_storage(x, st), to_storage(b, st)
    omega = 2.0 / 3.0
    got = oracle_jacobi(orc, g, active, xs, bs, omega, st)
    X, B = dense(g, from_storage(xs, st)), dense(g, from_storage(bs, st))
    S = neighbour_sum64(X)
    want = X + omega / 6.0 * (B + S - 6.0 * X)
    m = listed_mask(g, active)
    G = dense(g, from_storage(got, st))
    ulp = {0: 2.0 ** -23, 1: 2.0 ** -10, 2: 2.0 ** -7}[st]
    tol = 1e-6 * (np.abs(X) + omega / 6.0 * (np.abs(B) + np.abs(S) + 6 * np.abs(X))) + ulp * np.abs(want) + 1e-30
    assert np.all(np.abs(G - want)[m] <= tol[m])
    assert np.array_equal(G[~m], X[~m])  # only listed cells are written


def test_jacobi_constant_is_fixed_point(pkg, orc):
    g, active, x, b = band_problem(pkg, 40, 33, 48, 3, empty=0.75, full=0.75)
    x[:] = 0.75
    b[:] = 0.0
    for st in (0, 1, 2):
        xs = to_storage(x, st)
        assert np.array_equal(oracle_jacobi(orc, g, active, xs, to_storage(b, st), 2.0 / 3.0, st), xs)


@pytest.mark.parametrize("k", [(1, 0, 0), (2, 3, 1), (7, 5, 4)])
def test_jacobi_neumann_eigenmode_decay(pkg, orc, k):
    """With every block listed the mirrored operator has the eigenvectors
    prod cos(pi k_a (i_a + 1/2) / s_a), eigenvalue sum 2 (1 - cos(pi k_a / s_a));
    one Jacobi sweep with b = 0 scales a mode by 1 - omega lambda / 6."""
    s = (32, 48, 16)
    g = pkg.SparseGrid("eig", *s, 0.0, 0.0)
    g.ensure_blocks_parallel(np.arange(g.n_blocks))
    active = np.arange(g.n_blocks, dtype=np.int64)
    z, y, xx = np.meshgrid(*[np.arange(n) for n in s[::-1]], indexing="ij")
    mode = (np.cos(np.pi * k[0] * (xx + 0.5) / s[0]) * np.cos(np.pi * k[1] * (y + 0.5) / s[1])
            * np.cos(np.pi * k[2] * (z + 0.5) / s[2]))
    slab = np.zeros((g.n_slots, 4096), np.float32)
    bid = (xx >> 4) + (y >> 4) * g.nx + (z >> 4) * g.nxy
    vid = (xx & 15) | ((y & 15) << 4) | ((z & 15) << 8)
    slab[g.b2s[bid], vid] = mode
    lam = sum(2 * (1 - np.cos(np.pi * ka / sa)) for ka, sa in zip(k, s))
    omega = 2.0 / 3.0
    out = oracle_jacobi(orc, g, active, slab, np.zeros_like(slab), omega)
    got = dense(g, out)
    want = (1 - omega * lam / 6) * dense(g, slab)
    assert np.abs(got - want).max() < 2e-6


@pytest.mark.parametrize("st", [0, 1, 2])
def test_rbgs_equals_serial_gauss_seidel_in_any_order(pkg, orc, st):
    """One red-black half sweep pair = serial Gauss-Seidel over the red cells,
    then the black cells, each visited in a random order with exact float32
    arithmetic (same-colour cells never couple)."""
    g, active, x, b = band_problem(pkg, 20, 18, 17, 4, frac=0.8, n_full=1, n_unlisted=1)
    omega = np.float32(1.7)
    w6 = np.float32(omega / np.float32(6.0))
    xs, bs = to_storage(x, st), to_storage(b, st)
    got = oracle_rbgs(orc, g, active, xs, bs, 0, float(omega), st)
    got = oracle_rbgs(orc, g, active, got, bs, 1, float(omega), st)
    D = dense(g, from_storage(xs, st)).astype(np.float32)
    Bd = dense(g, from_storage(bs, st)).astype(np.float32)
    m = listed_mask(g, active)
    rng = np.random.default_rng(5)
    sz, sy, sx = D.shape
    cells = np.argwhere(m)
    for color in (0, 1):
        cc = cells[(cells.sum(1) & 1) == color]
        for z, y, x_ in cc[rng.permutation(len(cc))]:
            def at(zz, yy, xx):
                zz = min(max(zz, 0), sz - 1) if zz in (-1, sz) else zz
                yy = min(max(yy, 0), sy - 1) if yy in (-1, sy) else yy
                xx = min(max(xx, 0), sx - 1) if xx in (-1, sx) else xx
                return D[zz, yy, xx]
            s = addf(addf(addf(addf(addf(at(z, y, x_ - 1), at(z, y, x_ + 1)), at(z, y - 1, x_)), at(z, y + 1, x_)),
                                 at(z - 1, y, x_)), at(z + 1, y, x_))
            xc = D[z, y, x_]
            t = fmaf(np.float32(-6.0), xc, addf(Bd[z, y, x_], s))
            D[z, y, x_] = from_storage(to_storage(np.array([fmaf(w6, t, xc)], np.float32), st), st)[0]
    G = dense(g, from_storage(got, st)).astype(np.float32)
    assert np.array_equal(G.view(np.uint32), D.view(np.uint32))


def test_residual_and_fp64_norms(pkg, orc):
    g, active, x, b = band_problem(pkg, 40, 33, 48, 6)
    r, r2, b2 = oracle_residual(orc, g, active, x, b)
    X, B = dense(g, x), dense(g, b)
    S = neighbour_sum64(X)
    m = listed_mask(g, active)
    R = dense(g, r)
    want = B + S - 6 * X
    assert np.all(np.abs(R - want)[m] <= 1e-6 * (np.abs(B) + np.abs(S) + 6 * np.abs(X))[m])
    assert abs(r2 - np.sum(R[m] ** 2)) <= 1e-12 * r2
    assert abs(b2 - np.sum(B[m] ** 2)) <= 1e-12 * b2


def assemble(g, active):
    """Independent sparse FP64 matrix of the band operator with zero boundary
    values (absent / full / unlisted neighbours contribute nothing)."""
    m = listed_mask(g, active)
    ids = -np.ones(m.shape, np.int64)
    ids[m] = np.arange(m.sum())
    zz, yy, xx = np.nonzero(m)
    n = zz.size
    diag = np.full(n, 6.0)
    rows, cols = [np.arange(n)], [np.arange(n)]
    vals = [None]
    for dz, dy, dx in [(0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0), (-1, 0, 0), (1, 0, 0)]:
        z2, y2, x2 = zz + dz, yy + dy, xx + dx
        edge = (z2 < 0) | (z2 >= g.sz) | (y2 < 0) | (y2 >= g.sy) | (x2 < 0) | (x2 >= g.sx)
        diag[edge] -= 1.0  # the mirrored neighbour is the cell itself
        ok = ~edge
        j = ids[z2[ok], y2[ok], x2[ok]]
        inn = j >= 0
        rows.append(np.arange(n)[ok][inn])
        cols.append(j[inn])
    vals = [diag] + [-np.ones(r.size) for r in rows[1:]]
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n)), m


def test_cg64_matches_sparse_direct_solve(pkg, orc):
    g, active, x, b = band_problem(pkg, 32, 32, 24, 7, frac=0.6, empty=0.0, full=0.0)
    A, m = assemble(g, active)
    xd, it = oracle_cg64(orc, g, active, b.astype(np.float64), tol=1e-13)
    want = spl.spsolve(A.tocsc(), dense(g, b)[m])
    got = dense_f64(g, xd)[m]
    assert it > 0
    assert np.abs(got - want).max() <= 1e-9 * np.abs(want).max()


def dense_f64(g, slab64):
    return dense_from_slabs(g.b2s, slab64.reshape(-1, 4096), (g.sx, g.sy, g.sz))


@pytest.mark.parametrize("storage,refine", [(0, 0), (0, 1), (1, 1), (2, 1)])
def test_solve_reaches_tolerance_and_fp64_solution(pkg, orc, storage, refine):
    """Red-black SOR to the tolerance; plain FP32 sweeps stall near
    eps32 * ||A|| ||x|| / ||b|| (about 1.5e-6 on this thick band), refinement
    gets past it.  The result is compared with the FP64 CG solution."""
    g, active, x, b = band_problem(pkg, 48, 48, 48, 8, frac=0.5, n_full=0, n_unlisted=0, empty=0.0, full=0.0)
    tol = 1e-6 if refine else 5e-6
    xs, st = oracle_solve(orc, g, active, np.zeros_like(x), b, storage, refine, omega=1.95, per=10, max_sweeps=3000,
                          tol=tol)
    assert st["converged"] and st["rel_residual"] <= tol
    _, r2, b2 = oracle_residual(orc, g, active, xs, b)
    assert np.sqrt(r2 / b2) == pytest.approx(st["rel_residual"], rel=1e-12)
    xd, _ = oracle_cg64(orc, g, active, b.astype(np.float64), tol=1e-13)
    m = listed_mask(g, active)
    ref = dense_f64(g, xd)[m]
    err = np.abs(dense(g, xs)[m] - ref).max() / np.abs(ref).max()
    assert err < 10 * tol, err  # forward error <= cond * residual; cond ~ 10 here
