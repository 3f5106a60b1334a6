"""Oracle-side drivers for the transfer tests (test infrastructure)."""
import numpy as np


def oracle_p2g_mac(orc, s, pos, vel, mass, r_p, ref_commit_width=0):
    """particles_to_grid via the oracle: returns (active, b2s, slabs[6, n_slots, 4096])."""
    sx, sy, sz = s
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    vel = np.ascontiguousarray(vel, np.float32).reshape(-1, 3)
    mass = np.ascontiguousarray(mass, np.float32)
    n = pos.shape[0]
    nb = ((sx + 15) >> 4) * ((sy + 15) >> 4) * ((sz + 15) >> 4)
    bids = np.zeros(n, np.int64)
    orc.orc_particle_bids(sx, sy, sz, pos, n, bids)
    pb, starts, perm = orc.bucket_indices(bids, nb)
    active = orc.active_blocks(sx, sy, sz, pos, r_p + 1.0)
    b2s = np.zeros(nb, np.int32)
    b2s[active] = np.arange(2, 2 + active.size, dtype=np.int32)
    n_slots = 2 + active.size
    slabs = np.zeros((6, n_slots, 4096), np.float32)
    orc.orc_p2g_mac(sx, sy, sz, np.ascontiguousarray(pos[perm]), np.ascontiguousarray(vel[perm]),
                    np.ascontiguousarray(mass[perm]), pb, starts, pb.size, r_p, slabs, n_slots, b2s, active,
                    active.size, ref_commit_width)
    return active, b2s, slabs


def oracle_p2g_mac_f64(orc, s, pos, vel, mass, r_p):
    """The same splat summed in float64: returns (active, b2s, sums[6, n_slots,
    4096]) = [P_x, P_y, P_z, M_x, M_y, M_z], unnormalised."""
    sx, sy, sz = s
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    vel = np.ascontiguousarray(vel, np.float32).reshape(-1, 3)
    mass = np.ascontiguousarray(mass, np.float32)
    n = pos.shape[0]
    nb = ((sx + 15) >> 4) * ((sy + 15) >> 4) * ((sz + 15) >> 4)
    bids = np.zeros(n, np.int64)
    orc.orc_particle_bids(sx, sy, sz, pos, n, bids)
    pb, starts, perm = orc.bucket_indices(bids, nb)
    active = orc.active_blocks(sx, sy, sz, pos, r_p + 1.0)
    b2s = np.zeros(nb, np.int32)
    b2s[active] = np.arange(2, 2 + active.size, dtype=np.int32)
    n_slots = 2 + active.size
    sums = np.zeros((6, n_slots, 4096), np.float64)
    orc.orc_p2g_mac_f64(sx, sy, sz, np.ascontiguousarray(pos[perm]), np.ascontiguousarray(vel[perm]),
                        np.ascontiguousarray(mass[perm]), pb, starts, pb.size, r_p, sums, n_slots, b2s, active,
                        active.size)
    return active, b2s, sums


def dense_from_slabs(b2s, slab, s):
    """(sz, sy, sx) dense view of one channel."""
    sx, sy, sz = s
    nx, ny = (sx + 15) >> 4, (sy + 15) >> 4
    z, y, x = np.meshgrid(np.arange(sz), np.arange(sy), np.arange(sx), indexing="ij")
    bid = (x >> 4) + (y >> 4) * nx + (z >> 4) * nx * ny
    vid = (x & 15) | ((y & 15) << 4) | ((z & 15) << 8)
    return slab[b2s[bid], vid]


def oracle_p2g_bspline(orc, order, s, pos, vel, mass=None):
    sx, sy, sz = s
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    vel = np.ascontiguousarray(vel, np.float32).reshape(-1, 3)
    n = pos.shape[0]
    nb = ((sx + 15) >> 4) * ((sy + 15) >> 4) * ((sz + 15) >> 4)
    active = orc.active_blocks(sx, sy, sz, pos, 2.0)
    b2s = np.zeros(nb, np.int32)
    b2s[active] = np.arange(2, 2 + active.size, dtype=np.int32)
    n_slots = 2 + active.size
    slabs = np.zeros((4, n_slots, 4096), np.float32)
    m = None if mass is None else np.ascontiguousarray(mass, np.float32)
    orc.orc_p2g_bspline(order, sx, sy, sz, pos, vel, None if m is None else m.ctypes.data, n, b2s, n_slots, active,
                        active.size, slabs)
    return active, b2s, slabs


def oracle_g2p_bspline(orc, order, s, b2s, vel_slabs, pos):
    sx, sy, sz = s
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    out = np.zeros((pos.shape[0], 3), np.float32)
    vs = np.ascontiguousarray(vel_slabs, np.float32)
    orc.orc_g2p_bspline(order, sx, sy, sz, vs, vs.shape[1], b2s, pos, pos.shape[0], out)
    return out


def scatter_bspline_f64(orc, order, s, pos, vel, mass=None):
    """Brute-force per-particle scatter in float64 (dense mass, momentum)."""
    sx, sy, sz = s
    M = np.zeros((sz, sy, sx))
    P = np.zeros((3, sz, sy, sx))
    for i, p in enumerate(np.asarray(pos, np.float32)):
        f = np.floor(p).astype(int)
        m = 1.0 if mass is None else float(mass[i])
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    c = f + np.array([dx, dy, dz])
                    if (c < 0).any() or c[0] >= sx or c[1] >= sy or c[2] >= sz:
                        continue
                    w = 1.0
                    for a in range(3):
                        w *= orc.orc_bspline_weight(order, float(p[a]), int(c[a]))
                    M[c[2], c[1], c[0]] += w * m
                    P[:, c[2], c[1], c[0]] += w * m * np.asarray(vel[i], np.float64)
    return M, P
