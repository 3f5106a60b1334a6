"""Shared fixtures of the Poisson tests: band problems on small sparse grids,
an independent dense FP64 statement of the operator, and oracle wrappers."""
import ctypes as C
from fractions import Fraction

import numpy as np

from transfer_helpers import dense_from_slabs

STORAGE_DTYPE = {0: np.float32, 1: np.float16, 2: np.uint16}


def band_problem(pkg, sx, sy, sz, seed, frac=0.7, n_full=2, n_unlisted=2, empty=0.25, full=0.75):
    """A layout grid with value / absent / full blocks, the listed blocks (all
    value blocks but `n_unlisted`), and FP32 slabs x, b (slots 0 / 1 hold the
    absent / full neighbour values)."""
    rng = np.random.default_rng(seed)
    g = pkg.SparseGrid("band", sx, sy, sz, empty, full)
    bids = np.nonzero(rng.random(g.n_blocks) < frac)[0]
    g.ensure_blocks_parallel(bids)
    free = np.nonzero(g.b2s < 2)[0]
    rng.shuffle(free)
    for b in free[:n_full]:
        g.set_full_block(int(b))
    vb = np.nonzero(g.b2s >= 2)[0]
    rng.shuffle(vb)
    active = np.sort(vb[n_unlisted:]).astype(np.int64)
    n = g.n_slots
    x = rng.uniform(-1.0, 2.0, size=(n, 4096)).astype(np.float32)
    x[0], x[1] = empty, full
    b = rng.uniform(-1.0, 1.0, size=(n, 4096)).astype(np.float32)
    return g, active, x, b


def to_storage(a, st):
    a = np.ascontiguousarray(a, np.float32)
    if st == 0:
        return a.copy()
    if st == 1:
        return a.astype(np.float16)
    u = a.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def from_storage(a, st):
    if st == 0:
        return np.asarray(a, np.float32)
    if st == 1:
        return np.asarray(a, np.float16).astype(np.float32)
    return (np.asarray(a, np.uint16).astype(np.uint32) << 16).view(np.float32)


def listed_mask(g, active):
    """(sz, sy, sx) bool: cells of the listed blocks inside the domain."""
    m = np.zeros(g.n_blocks, bool)
    m[active] = True
    z, y, x = np.meshgrid(np.arange(g.sz), np.arange(g.sy), np.arange(g.sx), indexing="ij")
    return m[(x >> 4) + (y >> 4) * g.nx + (z >> 4) * g.nxy]


def dense(g, slab):
    return dense_from_slabs(g.b2s, np.asarray(slab).reshape(-1, 4096), (g.sx, g.sy, g.sz)).astype(np.float64)


def neighbour_sum64(X):
    """Sum of the 6 face neighbours with mirrored (Neumann) domain edges, FP64."""
    P = np.pad(X, 1, mode="symmetric")  # mirror: the cell past the edge is the edge cell
    return (P[1:-1, 1:-1, :-2] + P[1:-1, 1:-1, 2:] + P[1:-1, :-2, 1:-1] + P[1:-1, 2:, 1:-1] + P[:-2, 1:-1, 1:-1]
            + P[2:, 1:-1, 1:-1])


def oracle_residual(orc, g, active, x, b, st=0):
    r = np.zeros(x.shape, np.float32)
    r2, b2 = C.c_double(), C.c_double()
    orc.orc_poisson_residual(g.sx, g.sy, g.sz, st, x.ctypes.data, b.ctypes.data, g.b2s, active, active.size,
                             r.ctypes.data, C.byref(r2), C.byref(b2))
    return r, r2.value, b2.value


def oracle_jacobi(orc, g, active, x, b, omega, st=0):
    out = x.copy()
    orc.orc_jacobi(g.sx, g.sy, g.sz, st, x.ctypes.data, out.ctypes.data, b.ctypes.data, g.b2s, active,
                   active.size, omega)
    return out


def oracle_rbgs(orc, g, active, x, b, color, omega, st=0):
    out = x.copy()
    orc.orc_rbgs(g.sx, g.sy, g.sz, st, out.ctypes.data, b.ctypes.data, g.b2s, active, active.size, color, omega)
    return out


def oracle_solve(orc, g, active, x, b, storage=0, refine=0, omega=1.9, per=10, max_sweeps=1000, tol=1e-6):
    x = np.ascontiguousarray(x, np.float32).copy()
    out = np.zeros(4)
    orc.orc_poisson_solve(g.sx, g.sy, g.sz, x.reshape(-1), np.ascontiguousarray(b, np.float32).reshape(-1),
                          g.n_slots, g.b2s, active, active.size, storage, refine, omega, per, max_sweeps, tol, out)
    return x, {"sweeps": int(out[0]), "checks": int(out[1]), "converged": int(out[2]), "rel_residual": out[3]}


def oracle_cg64(orc, g, active, b, tol=1e-12, max_iter=20000):
    x = np.zeros(g.n_slots * 4096, np.float64)
    bb = np.ascontiguousarray(b, np.float64).reshape(-1)
    it = orc.orc_cg64(g.sx, g.sy, g.sz, x, bb, g.n_slots, g.b2s, active, active.size, tol, max_iter)
    return x.reshape(-1, 4096), it


# -- exact float32 arithmetic for the serial references ------------------------
def round_f32(e: Fraction) -> np.float32:
    """Correctly rounded (RNE) float32 of an exact rational."""
    f = np.float32(float(e))  # double rounding is repaired below
    fe = Fraction(float(f))
    if fe == e or not np.isfinite(f):
        return f
    g = np.nextafter(f, np.float32(np.inf) if e > fe else np.float32(-np.inf), dtype=np.float32)
    dg, df = abs(Fraction(float(g)) - e), abs(fe - e)
    if dg < df:
        return g
    if dg > df:
        return f
    return f if (int(np.float32(f).view(np.uint32)) & 1) == 0 else g


def fmaf(a, b, c) -> np.float32:
    return round_f32(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))


def addf(a, b) -> np.float32:
    return np.float32(np.float32(a) + np.float32(b))
