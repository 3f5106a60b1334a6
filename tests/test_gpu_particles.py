"""GPU parity of the particle path: Morton keys, radix sort, bucketing and
active-block compaction against the oracle (all bit-exact: integer work)."""
import numpy as np
import pytest

from msbg_rs_amd import particles as P

pytestmark = pytest.mark.gpu


def oracle_keys(orc, pos):
    pos = np.ascontiguousarray(pos, np.float32)
    keys = np.zeros(pos.shape[0], np.uint32)
    orc.orc_morton_keys(pos, pos.shape[0], keys)
    return keys


def oracle_perm(orc, keys):
    perm = np.zeros(keys.size, np.uint32)
    orc.orc_stable_sort_perm(np.ascontiguousarray(keys, np.uint32), keys.size, perm)
    return perm


def test_morton_keys_match_oracle(hip, orc):
    rng = np.random.default_rng(0)
    pos = rng.uniform(-2.0, 1100.0, size=(100_003, 3)).astype(np.float32)
    pos[:5] = [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1023.5, 1023.5, 1023.5]]
    got = P.morton_keys(pos, lib=hip).get()
    assert np.array_equal(got, oracle_keys(orc, pos))
    assert list(got[:4]) == [0, 1, 2, 4] and got[4] == (1 << 30) - 1


@pytest.mark.parametrize("n", [0, 1, 5, 4095, 4096, 4097, 65537, 1_000_003])
@pytest.mark.parametrize("bits", [30, 32, 12])
def test_radix_sort_is_stable_sort(hip, orc, n, bits):
    rng = np.random.default_rng(n + bits)
    hi = (1 << bits) - 1
    keys = rng.integers(0, hi + 1, size=n, dtype=np.uint64).astype(np.uint32)
    if n > 10:
        keys[::7] = keys[1]  # many duplicates
    vals = np.arange(n, dtype=np.uint32)
    dk, dv = hip.to_device(keys), hip.to_device(vals)
    P.radix_sort_pairs(dk, dv, bits, lib=hip)
    want = oracle_perm(orc, keys)
    assert np.array_equal(dv.get(), want)
    assert np.array_equal(dk.get(), keys[want])


def test_morton_sort_matches_oracle(hip, orc):
    rng = np.random.default_rng(3)
    pos = rng.uniform(0, 512, size=(300_000, 3)).astype(np.float32)
    pos[100:200] = pos[50]  # equal keys keep input order
    keys, perm = P.morton_sort(pos, lib=hip)
    k = oracle_keys(orc, pos)
    want = oracle_perm(orc, k)
    assert np.array_equal(perm.get(), want) and np.array_equal(keys.get(), k[want])


@pytest.mark.parametrize("n,with_mass", [(0, True), (1, True), (4097, False), (300_000, True)])
def test_morton_sort_particles_matches_sort_and_gathers(hip, orc, n, with_mass):
    """The fused sort: keys and permutation as the oracle's stable Morton
    sort, positions / velocities / masses the rows of that permutation (bit
    for bit; masses 1 when none are given)."""
    rng = np.random.default_rng(n + 17)
    pos = rng.uniform(0, 700, size=(n, 3)).astype(np.float32)
    if n > 200:
        pos[100:200] = pos[50]  # equal keys keep input order
    vel = rng.normal(size=(n, 3)).astype(np.float32)
    mass = rng.uniform(0.5, 2.0, n).astype(np.float32) if with_mass else None
    keys, perm, ps, vs, ms = P.morton_sort_particles(pos, vel, mass, lib=hip)
    k = oracle_keys(orc, pos)
    want = oracle_perm(orc, k)
    assert np.array_equal(perm.get(), want) and np.array_equal(keys.get(), k[want])
    assert np.array_equal(ps.get().reshape(n, 3), pos[want])
    assert np.array_equal(vs.get().reshape(n, 3), vel[want])
    assert np.array_equal(ms.get(), mass[want] if with_mass else np.ones(n, np.float32))


@pytest.mark.parametrize("n,nb", [(0, 1), (1, 1), (4, 2), (1000, 64), (10_000, 1024), (1_000_000, 32768),
                                  (200_000, 262144)])
def test_bucket_indices_match_oracle(hip, orc, n, nb):
    rng = np.random.default_rng(n)
    bids = rng.integers(0, nb, n).astype(np.int64)
    if n == 4:
        bids = np.array([1, 0, 1, 0])  # bucket_01
    plan = P.bucket_indices(bids, nb, lib=hip)
    pb, st, perm = plan.host()
    wpb, wst, wperm = orc.bucket_indices(bids, nb)
    assert np.array_equal(pb, wpb) and np.array_equal(st, wst) and np.array_equal(perm, wperm)


def test_particle_bids_match_oracle(hip, orc):
    rng = np.random.default_rng(4)
    pos = rng.uniform(0, [80, 48, 33], size=(50_000, 3)).astype(np.float32)
    got = P.particle_bids(80, 48, 33, pos, lib=hip).get()
    want = np.zeros(pos.shape[0], np.int64)
    orc.orc_particle_bids(80, 48, 33, pos, pos.shape[0], want)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("shape,r_scan,n", [((64, 64, 64), 4.0, 1), ((128, 128, 128), 3.0, 5000),
                                            ((512, 512, 512), 3.0, 200_000), ((32, 48, 16), 4.0, 3)])
def test_active_blocks_match_oracle(hip, orc, shape, r_scan, n):
    rng = np.random.default_rng(n)
    pos = rng.uniform(0, shape, size=(n, 3)).astype(np.float32)
    if n == 3:
        pos = np.array([[17.0, 33.0, 8.0], [0.1, 0.1, 0.1], [31.9, 47.9, 15.9]], np.float32)
    got = P.active_blocks(*shape, pos, r_scan, lib=hip)
    assert np.array_equal(got, orc.active_blocks(*shape, pos, r_scan))


def test_active_blocks_known_answers_on_device(hip):
    assert list(P.active_blocks(64, 64, 64, [[15.5, 8.0, 8.0]], 4.0, lib=hip)) == [0, 1]
    assert list(P.active_blocks(32, 48, 16, [[17.0, 33.0, 8.0]], 4.0, lib=hip)) == [2, 3, 4, 5]
    assert P.active_blocks(64, 64, 64, np.zeros((0, 3), np.float32), 4.0, lib=hip).size == 0


def test_gather_rows(hip):
    rng = np.random.default_rng(5)
    src = rng.normal(size=(1000, 3)).astype(np.float32)
    perm = rng.permutation(1000).astype(np.uint32)
    got = P.gather_rows(hip.to_device(src), hip.to_device(perm), lib=hip).get()
    assert np.array_equal(got, src[perm])
