"""Pins the oracle's particle functions on the reference's known answers
(particles/sort.rs:352-451, particles/active.rs:72-187) and on closed forms."""
import numpy as np
import pytest


def test_bucket_known_answers(orc):
    # bucket_01 (sort.rs:353-366): bids [1, 0, 1, 0]
    pb, st, perm = orc.bucket_indices([1, 0, 1, 0], 2)
    assert list(pb) == [0, 1] and list(st) == [0, 2, 4]
    assert sorted(perm[0:2]) == [1, 3] and sorted(perm[2:4]) == [0, 2]
    assert list(perm) == [1, 3, 0, 2]  # stable (the dense oracle's order)
    # bucket_02 / 03 / 04
    pb, st, perm = orc.bucket_indices([3], 8)
    assert list(pb) == [3] and list(st) == [0, 1]
    pb, st, perm = orc.bucket_indices([0, 0, 0], 1)
    assert list(pb) == [0] and list(st) == [0, 3] and list(perm) == [0, 1, 2]
    pb, st, perm = orc.bucket_indices([], 1)
    assert pb.size == 0 and list(st) == [0] and perm.size == 0


def test_bucket_matches_stable_argsort(orc):
    """Closed form: the dense counting sort is numpy's stable argsort by bid."""
    rng = np.random.default_rng(1)
    for n, nb in [(1000, 64), (10000, 1024), (1, 1)]:
        bids = rng.integers(0, nb, n)
        pb, st, perm = orc.bucket_indices(bids, nb)
        assert np.array_equal(perm, np.argsort(bids, kind="stable"))
        u, c = np.unique(bids, return_counts=True)
        assert np.array_equal(pb, u) and np.array_equal(st, np.concatenate([[0], np.cumsum(c)]))


def test_active_blocks_known_answers(orc):
    r_scan = 4.0  # r_particle 2 + nb_dist 2 (active.rs:57-69)
    assert list(orc.active_blocks(64, 64, 64, [[8.0, 8.0, 8.0]], r_scan)) == [0]  # active_01
    assert list(orc.active_blocks(64, 64, 64, [[15.5, 8.0, 8.0]], r_scan)) == [0, 1]  # active_02
    assert list(orc.active_blocks(64, 64, 64, [[0.1, 0.1, 0.1]], r_scan)) == [0]  # active_03
    a = orc.active_blocks(64, 64, 64, [[63.0, 63.0, 63.0]], r_scan)  # active_04
    assert (a[0] % 4, (a[0] // 4) % 4, a[0] // 16) == (3, 3, 3)
    assert orc.active_blocks(64, 64, 64, np.zeros((0, 3)), r_scan).size == 0  # active_05
    assert list(orc.active_blocks(64, 64, 64, [[8.0, 8.0, 8.0], [9.0, 9.0, 9.0]], r_scan)) == [0]  # active_06
    assert list(orc.active_blocks(32, 48, 16, [[17.0, 33.0, 8.0]], r_scan)) == [2, 3, 4, 5]  # active_07


def test_active_blocks_dense_reference(orc):
    """active_10 (active.rs:171-208): the union equals a brute-force footprint set."""
    s = np.uint32(0x12345678)
    pos = []
    for _ in range(5000):
        row = []
        for _ in range(3):
            s = np.uint32((int(s) * 1664525 + 1013904223) & 0xFFFFFFFF)
            row.append(float((int(s) >> 8) % 128))
        pos.append(row)
    pos = np.array(pos, np.float32)
    want = set()
    rb = np.float32(4.0) / np.float32(16.0)
    for p in pos:
        lo = [max(int(np.trunc(np.float32(c) / np.float32(16.0) - rb)), 0) for c in p]
        hi = [min(int(np.trunc(np.float32(c) / np.float32(16.0) + rb)), 7) for c in p]
        for bz in range(lo[2], hi[2] + 1):
            for by in range(lo[1], hi[1] + 1):
                for bx in range(lo[0], hi[0] + 1):
                    want.add(bx + by * 8 + bz * 64)
    assert list(orc.active_blocks(128, 128, 128, pos, 4.0)) == sorted(want)


def test_morton_keys_and_stable_perm(orc, pkg):
    rng = np.random.default_rng(2)
    pos = rng.uniform(0, 1023.99, size=(2000, 3)).astype(np.float32)
    keys = np.zeros(2000, np.uint32)
    orc.orc_morton_keys(pos, 2000, keys)
    for i in range(0, 2000, 97):
        x, y, z = (int(c) for c in pos[i])
        assert keys[i] == pkg.morton3(x, y, z) & 0x3FFFFFFF
    keys[::3] = keys[0]  # duplicates exercise stability
    perm = np.zeros(2000, np.uint32)
    orc.orc_stable_sort_perm(keys, 2000, perm)
    assert np.array_equal(perm, np.argsort(keys, kind="stable"))
