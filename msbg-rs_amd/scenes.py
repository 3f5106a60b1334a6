"""Seeded synthetic scenes for the benchmark workloads and the parity tests.

Random numbers come from the LCG the reference benches use
(`s = s*1664525 + 1013904223 mod 2^32`, `u = s / u32::MAX`,
benches/velocity_bench.rs:45-60), generated with exact jump-ahead so a
stream of millions of values is vectorized yet identical to the sequential
one.
"""
import numpy as np

from .grid import SparseGrid

A, C, M32 = 1664525, 1013904223, 0xFFFFFFFF


def lcg_u32(seed: int, n: int) -> np.ndarray:
    """The first n states s_1..s_n of the LCG started at `seed` (uint32)."""
    n = int(n)
    if n == 0:
        return np.zeros(0, np.uint32)
    m = min(n, 4096)
    # s_{j} = Aj * s_0 + Cj for j = 1..m
    Aj = np.zeros(m, np.uint64)
    Cj = np.zeros(m, np.uint64)
    a, c = 1, 0
    for j in range(m):
        a = (a * A) & M32
        c = (c * A + C) & M32
        Aj[j], Cj[j] = a, c
    am, cm = int(Aj[-1]), int(Cj[-1])
    nb = (n + m - 1) // m
    starts = np.zeros(nb, np.uint64)
    s = int(seed) & M32
    for b in range(nb):
        starts[b] = s
        s = (am * s + cm) & M32
    out = (starts[:, None] * Aj[None, :] + Cj[None, :]) & np.uint64(M32)
    return out.ravel()[:n].astype(np.uint32)


def lcg_uniform(seed: int, n: int) -> np.ndarray:
    """`seed as f32 / u32::MAX as f32` per draw (float32 in [0, 1])."""
    return (lcg_u32(seed, n).astype(np.float32) / np.float32(4294967295.0)).astype(np.float32)


def lcg_normal(seed: int, n: int) -> np.ndarray:
    """Box-Muller normals from the LCG stream (float32)."""
    k = (n + 1) // 2
    u = lcg_uniform(seed, 2 * k).astype(np.float64)
    u1 = np.clip(u[0::2], 1e-12, 1.0)
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
    return z[:n].astype(np.float32)


# ---------------------------------------------------------------------------
# Config 1: 128^3 spherical shell, one level.
# ---------------------------------------------------------------------------
def shell_scene(S=128, r_in=40.0, r_out=56.0, r_mid=48.0, seed=1):
    """Active 16^3 blocks intersecting the shell r in [r_in, r_out] around a
    jittered centre, phi = tanh((r - r_mid)/2) at cell centres."""
    j = lcg_uniform(seed, 3) - np.float32(0.5)
    c = np.array([S / 2, S / 2, S / 2], np.float64) + j.astype(np.float64)
    g = SparseGrid("phi", S, S, S, 0.0, 1.0)
    bx, by, bz = np.meshgrid(np.arange(g.nx), np.arange(g.ny), np.arange(g.nz), indexing="ij")
    lo = np.stack([bx, by, bz], -1).reshape(-1, 3) * 16.0
    hi = lo + 16.0
    nearest = np.clip(c, lo, hi)
    dmin = np.linalg.norm(nearest - c, axis=1)
    far = np.maximum(np.abs(lo - c), np.abs(hi - c))
    dmax = np.linalg.norm(far, axis=1)
    hit = (dmin <= r_out) & (dmax >= r_in)
    coords = np.stack([bx, by, bz], -1).reshape(-1, 3)[hit]
    bids = np.sort(coords[:, 0] + coords[:, 1] * g.nx + coords[:, 2] * g.nxy).astype(np.int64)

    def phi(X, Y, Z):
        r = np.sqrt((X + 0.5 - c[0]) ** 2 + (Y + 0.5 - c[1]) ** 2 + (Z + 0.5 - c[2]) ** 2)
        return np.tanh((r - r_mid) / 2.0).astype(np.float32)

    g.fill_from_function(bids, phi)
    return g, bids, c


def shell_particles(n, c, r_in=40.0, r_out=56.0, seed_pos=2, seed_vel=3):
    """n particles uniform inside the shell (rejection on the LCG stream) with
    N(0,1) float32 velocities."""
    out = []
    have = 0
    batch = max(4 * n, 1024)
    s = seed_pos
    while have < n:
        u = lcg_uniform(s, 3 * batch).reshape(-1, 3).astype(np.float64)
        s = int(lcg_u32(s, 3 * batch)[-1])
        p = c + (u * 2.0 - 1.0) * r_out
        r = np.linalg.norm(p - c, axis=1)
        keep = p[(r >= r_in) & (r <= r_out)]
        out.append(keep)
        have += keep.shape[0]
    pos = np.concatenate(out)[:n].astype(np.float32)
    vel = lcg_normal(seed_vel, 3 * n).reshape(n, 3)
    return pos, vel


# ---------------------------------------------------------------------------
# Config 2: 512^3 noisy sphere, 4 levels.
# ---------------------------------------------------------------------------
def value_noise(seed, lattice, period):
    """Smooth value noise: LCG values on a lattice of `period`-cell spacing,
    smoothstep-trilinear between lattice points.  Returns f(X, Y, Z) in [-1, 1]."""
    vals = (lcg_uniform(seed, lattice ** 3).reshape(lattice, lattice, lattice) * 2.0 - 1.0).astype(np.float64)

    def f(X, Y, Z):
        out = np.zeros(np.broadcast(X, Y, Z).shape)
        q = [np.asarray(v, np.float64) / period for v in (X, Y, Z)]
        i0 = [np.clip(np.floor(v).astype(np.int64), 0, lattice - 2) for v in q]
        t = [np.clip(v - i, 0.0, 1.0) for v, i in zip(q, i0)]
        t = [w * w * (3 - 2 * w) for w in t]
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    w = (t[0] if dx else 1 - t[0]) * (t[1] if dy else 1 - t[1]) * (t[2] if dz else 1 - t[2])
                    out += w * vals[i0[0] + dx, i0[1] + dy, i0[2] + dz]
        return out

    return f


class MultiresScene:
    """A multires grid: per-level host SparseGrids (block edge 16 >> l) over one
    block grid, a per-block level map, and the signed distance used to fill it."""

    def __init__(self, S, n_levels, levels, grids, sdf, centre, radius):
        self.S, self.n_levels, self.levels, self.grids = S, n_levels, levels, grids
        self.sdf, self.centre, self.radius = sdf, centre, radius

    def level_blocks(self, l):
        return np.nonzero(self.levels == l)[0].astype(np.int64)


def noisy_sphere_scene(S=512, n_levels=4, band=8.0, seed=4, amp=6.0, regularize=None):
    """Noisy sphere R = 0.35*S.  Level 0 = blocks within `band` cells of the
    surface, coarser levels outward by distance bands, 2:1-regularized with
    `regularize(levels, nx, ny, nz, n_levels)` (refinement.rs:66-107) if given.
    Each level l holds its blocks' phase field phi = tanh(sdf/2) sampled at
    its own cell centres."""
    nb = S // 16
    R = 0.35 * S
    c = np.array([S / 2.0] * 3) + (lcg_uniform(seed + 100, 3).astype(np.float64) - 0.5)
    noise = value_noise(seed, 9, S / 8.0)

    def sdf(X, Y, Z):
        d = np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2)
        return d - R - amp * noise(X, Y, Z)

    b = np.arange(nb)
    BZ, BY, BX = np.meshgrid(b, b, b, indexing="ij")
    ctr = [(v.ravel() + 0.5) * 16.0 for v in (BX, BY, BZ)]
    d = np.abs(sdf(*ctr))
    half_diag = 8.0 * np.sqrt(3.0)
    reach = d - half_diag  # lower bound of |sdf| inside the block (the noise is smooth)
    levels = np.full(nb ** 3, n_levels - 1, np.uint8)
    bands = [band] + [band + 32.0 * (2 ** k - 1) + 16.0 for k in range(1, n_levels - 1)]
    for l in range(n_levels - 2, -1, -1):
        levels[reach <= bands[l]] = l
    if regularize is not None:
        regularize(levels, nb, nb, nb, n_levels)
    grids = []
    for l in range(n_levels):
        L = 4 - l
        g = SparseGrid("phi_l%d" % l, S >> l, S >> l, S >> l, 0.0, 0.0, log2_bsx=L)
        bids = np.nonzero(levels == l)[0].astype(np.int64)
        scale = float(1 << l)

        def phi(X, Y, Z, scale=scale):
            return np.tanh(sdf((X + 0.5) * scale, (Y + 0.5) * scale, (Z + 0.5) * scale) / 2.0).astype(np.float32)

        if bids.size:
            for k in range(0, bids.size, 2048):
                g.fill_from_function(bids[k:k + 2048], phi)
        grids.append(g)
    return MultiresScene(S, n_levels, levels, grids, sdf, c, R)


def band_particles(scene: "MultiresScene", n, seed_pos=5, seed_vel=6):
    """n particles uniform over the level-0 blocks of a multires scene (each
    particle picks a block, then a position inside it) with N(0,1) float32
    velocities; positions stay strictly inside the domain."""
    blocks = scene.level_blocks(0)
    nb = scene.S // 16
    u = lcg_uniform(seed_pos, 4 * n).reshape(n, 4).astype(np.float64)
    pick = blocks[np.minimum((u[:, 0] * blocks.size).astype(np.int64), blocks.size - 1)]
    org = np.stack([pick % nb, (pick // nb) % nb, pick // (nb * nb)], 1) * 16.0
    pos = (org + np.minimum(u[:, 1:] * 16.0, 16.0 - 1e-3)).astype(np.float32)
    vel = lcg_normal(seed_vel, 3 * n).reshape(n, 3)
    return pos, vel


def reference_shell_blocks(nx, ny, nz):
    """generate_active_blocks (benches/allocator_benches.rs:233-256): blocks whose
    centre lies in the shell [0.9, 1.0] * min(n)/2."""
    cx, cy, cz = nx * 0.5, ny * 0.5, nz * 0.5
    r_out = min(nx, ny, nz) * 0.5
    r_in = r_out * 0.9
    BZ, BY, BX = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    d = np.sqrt((BX + 0.5 - cx) ** 2 + (BY + 0.5 - cy) ** 2 + (BZ + 0.5 - cz) ** 2).ravel()
    sel = (d >= r_in) & (d <= r_out)
    return np.nonzero(sel)[0].astype(np.int64)


def poisson_rhs(grid: SparseGrid, active, seed=11):
    """Right-hand side slab (n_slots x 4096, float32) of the band Poisson
    problems: LCG-uniform values in [-1, 1] on the cells of the listed blocks
    (block k of `active` takes draws k*4096 .. k*4096+4095), zero elsewhere."""
    active = np.asarray(active, np.int64)
    b = np.zeros((grid.n_slots, grid.N), np.float32)
    u = lcg_uniform(seed, active.size * grid.N).reshape(active.size, grid.N)
    b[grid.b2s[active]] = (u * np.float32(2.0) - np.float32(1.0)).astype(np.float32)
    return b


# ---------------------------------------------------------------------------
# Config 5: dam-break box + droplets (the FLIP step's particles).
def dam_break(S, n, seed=21, box=(0.35, 0.6, 0.5), n_drops=6, drop_frac=0.1):
    """n particles: (1 - drop_frac) n uniform in the box [0, box * S) at the
    domain's low corner (at rest plus N(0, 0.1) jitter), the rest in n_drops
    spheres of radius S/24 above it falling at -1 (plus the same jitter);
    unit masses.  Positions lie in [0.5, S - 0.5) per axis."""
    n_drop = int(n * drop_frac)
    n_box = n - n_drop
    u = lcg_uniform(seed, 3 * n_box).reshape(n_box, 3).astype(np.float64)
    pos_box = 0.5 + u * (np.array(box) * S - 0.5)
    c = lcg_uniform(seed + 1, 3 * n_drops).reshape(n_drops, 3).astype(np.float64)
    centres = np.stack([S * (0.45 + 0.4 * c[:, 0]), S * (0.7 + 0.2 * c[:, 1]), S * (0.15 + 0.7 * c[:, 2])], 1)
    r = S / 24.0
    v = lcg_normal(seed + 2, 3 * n_drop).reshape(n_drop, 3).astype(np.float64)
    v /= np.maximum(np.linalg.norm(v, axis=1, keepdims=True), 1e-12)
    rad = r * lcg_uniform(seed + 3, n_drop).astype(np.float64) ** (1.0 / 3.0)
    pick = np.arange(n_drop) % max(n_drops, 1)
    pos_drop = centres[pick] + v * rad[:, None]
    pos = np.clip(np.concatenate([pos_box, pos_drop]), 0.5, S - 0.5 - 1e-3).astype(np.float32)
    vel = (0.1 * lcg_normal(seed + 4, 3 * n).reshape(n, 3)).astype(np.float32)
    vel[n_box:, 1] -= np.float32(1.0)
    mass = np.ones(n, np.float32)
    return pos, vel, mass
