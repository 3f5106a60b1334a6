"""Pins the CPU oracle against the reference's own golden vectors.

Every expected value here comes from the reference's difftests
(tests/golden/reference_goldens.json, produced there by the C++ MSBG
baseline) or from one of the reference's known-answer unit tests.  Inputs are
the same closed-form fields the reference tests use, evaluated in float32 in
the same operation order.
"""
import json
import os
import struct

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_goldens.json")))
f32 = np.float32


def fnv1a(h, data: bytes):
    for b in data:
        h ^= b
        h = (h * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def cpp_script_grid(pkg):
    """tests/difftest_cpp.rs:22-35: the fixed write script over a 17x33x5 grid."""
    g = pkg.SparseGrid("diff", 17, 33, 5, 0.0, 1.0)
    g.set_voxel(0, 0, 0, 42.0)
    g.set_voxel(16, 0, 0, 7.0)
    g.set_voxel(0, 32, 0, 3.0)
    g.set_voxel(16, 32, 4, 99.0)
    g.set_voxel(5, 5, 2, 123.5)
    g.set_voxel(3, 3, 3, 0.25)
    g.set_full_block(g.get_block_id(16, 16, 0))
    g.set_empty_block(g.get_block_id(0, 16, 0))
    return g


def grid_hash(dense):
    """FNV-1a over every voxel's f32 bits, z, y, x order (difftest_cpp.rs:37-45)."""
    return fnv1a(0xCBF29CE484222325, np.ascontiguousarray(dense, dtype="<f4").tobytes())


def test_difftest_cpp_hash(pkg):
    g = cpp_script_grid(pkg)
    h = grid_hash(g.to_dense())
    assert h == int(GOLD["difftest_cpp"]["hash"], 16)
    # The hash sees every voxel: a one-voxel change must break it.
    g2 = cpp_script_grid(pkg)
    g2.set_voxel(5, 5, 2, 123.0)
    assert grid_hash(g2.to_dense()) != h


def stencil_field(X, Y, Z):
    """tests/difftest_stencil.rs:25-32 in float32 order."""
    u = (X - 24).astype(f32)
    v = (Y - 24).astype(f32)
    w = (Z - 24).astype(f32)
    q = f32(0.01) * (u * u + f32(2.0) * v * v + f32(3.0) * w * w + u * v)
    r = f32(0.001) * (u * u * u * u + v * v * v * v + w * w * w * w)
    return q + r


def build_stencil_grid(pkg):
    g = pkg.SparseGrid("meancurvtest", 64, 64, 64, 0.0, 1.0)
    g.fill_from_function(np.arange(g.n_blocks), stencil_field)
    return g


@pytest.mark.parametrize("name,stencil,halo,full", [("laplacian", 0, 1, False), ("meancurv", 1, 1, True),
                                                    ("bilaplacian", 2, 2, True)])
def test_difftest_stencil_goldens(pkg, orc, name, stencil, halo, full):
    d = GOLD["difftest_stencil"]
    g = build_stencil_grid(pkg)
    bid = g.get_block_id(16, 16, 16)
    tile = orc.halo_fill(g, bid, halo, full, bc=1)
    out = orc.kernel(stencil, tile, d["dt"])
    got = out[d["sample_idx"]]
    np.testing.assert_allclose(got, d["golden"][name], atol=d["tol"], rtol=0)
    # The restatement sits within 1e-5 of the C++ baseline (measured <= 6e-6),
    # far inside the reference's 1e-4: a dropped or mis-signed tap moves these
    # samples by >= 1e-3.
    np.testing.assert_allclose(got, d["golden"][name], atol=1e-5, rtol=0)


def smoother_field(X, Y, Z):
    """tests/difftest_smoother.rs:189-191."""
    return np.sin(f32(0.1) * X.astype(f32)) * np.cos(f32(0.08) * Y.astype(f32)) * np.sin(f32(0.06) * Z.astype(f32))


@pytest.mark.parametrize("name,stencil", [("laplacian", 0), ("meancurv", 1)])
def test_difftest_smoother_goldens(pkg, orc, name, stencil):
    d = GOLD["difftest_smoother"]
    g = pkg.SparseGrid("smoothertest", 64, 64, 64, 0.0, 1.0)
    g.fill_from_function(np.arange(g.n_blocks), smoother_field)
    active = np.arange(g.n_blocks)
    orc.sweep(g, active, stencil, d["dt"], d["iterations"], nthreads=4)
    flat = np.concatenate([g.block_data(b) for b in range(g.n_blocks)])
    np.testing.assert_allclose(flat[d["sample_idx"]], d["golden"][name], atol=d["tol"][name], rtol=0)
    np.testing.assert_allclose(flat[d["sample_idx"]], d["golden"][name], atol=1e-6, rtol=0)


def interp_field(X, Y, Z):
    """tests/difftest_interp.rs:89-93."""
    x, y, z = X.astype(f32), Y.astype(f32), Z.astype(f32)
    return (f32(0.002) * x * x + f32(0.003) * y * y + f32(0.004) * z * z + f32(0.005) * x * y
            + f32(0.006) * x * z + f32(0.007) * y * z + f32(0.1) * x + f32(0.05) * y - f32(0.02) * z + f32(0.75))


def test_difftest_interp_goldens(pkg, orc):
    d = GOLD["difftest_interp"]
    g = pkg.SparseGrid("interp", 48, 48, 48, 0.0, 1.0)
    g.fill_from_function(np.arange(g.n_blocks), interp_field)
    pos = np.array(d["positions"], np.float32)
    CORNER, CLAMP = 0, 0
    lv, lg, _ = orc.sample(g, pos, 0, CORNER, CLAMP, want_grad=True)
    cv, cg, ch = orc.sample(g, pos, 1, CORNER, CLAMP, want_grad=True, want_hess=True)
    got = np.concatenate([lv[:, None], lg, cv[:, None], cg, ch], axis=1)
    np.testing.assert_allclose(got, np.array(d["golden"], np.float32), atol=d["tol"], rtol=0)
    np.testing.assert_allclose(got, np.array(d["golden"], np.float32), atol=1e-5, rtol=0)


# ---------------------------------------------------------------------------
# Known-answer unit tests of the reference (halo.rs:433-722, sort.rs:63-100).
# ---------------------------------------------------------------------------

def bid_grid(pkg, s=48):
    g = pkg.SparseGrid("halo_test", s, s, s, 0.0, 1.0)
    bids = np.arange(g.n_blocks)
    g.ensure_blocks_parallel(bids)
    for b in bids:
        g.block_data(b)[:] = b
    return g


def test_h04_all_neighbors_full_fill(pkg, orc):
    g = bid_grid(pkg)
    t = orc.halo_fill(g, 13, 1, True)  # t[z, y, x]
    at = lambda x, y, z: t[z, y, x]
    assert at(1, 1, 1) == 13 and at(0, 1, 1) == 12 and at(17, 1, 1) == 14
    assert at(1, 0, 1) == 10 and at(1, 17, 1) == 16 and at(1, 1, 0) == 4 and at(1, 1, 17) == 22
    assert at(0, 0, 0) == 0 and at(17, 17, 17) == 26 and at(0, 0, 1) == 9 and at(0, 17, 17) == 24


def test_h05_h07_dirichlet_uses_empty(pkg, orc):
    g = pkg.SparseGrid("h", 16, 16, 16, 0.0, 1.0)
    g.set_voxel(0, 0, 0, 42.0)
    g.set_empty_value(7.0)
    t = orc.halo_fill(g, 0, 1, True, bc=2)
    assert np.all(t[1:17, 1:17, 0] == 7.0)
    g = pkg.SparseGrid("h", 48, 48, 48, 0.0, 1.0)
    g.set_voxel(0, 0, 0, 5.0)
    g.set_empty_value(2.0)
    t = orc.halo_fill(g, 0, 1, True, bc=2)
    assert t[1, 1, 1] == 5.0 and t[1, 1, 0] == 2.0 and t[1, 0, 1] == 2.0 and t[0, 1, 1] == 2.0
    assert t[1, 1, 17] == 2.0 and t[0, 0, 0] == 2.0


def axis_grid(pkg, sx, sy, sz):
    g = pkg.SparseGrid("h", sx, sy, sz, 0.0, 1.0)
    g.fill_from_function(np.arange(g.n_blocks), lambda X, Y, Z: X.astype(np.float32))
    return g


def test_h06_h08_neumann_mirror_and_partial_block(pkg, orc):
    t = orc.halo_fill(axis_grid(pkg, 16, 16, 16), 0, 1, True)
    assert t[1, 1, 0] == 0.0 and t[1, 1, 17] == 15.0
    g = axis_grid(pkg, 17, 16, 16)
    t = orc.halo_fill(g, g.get_block_id(16, 0, 0), 1, True)
    assert t[1, 1, 1] == 16.0 and t[1, 1, 2] == 16.0 and t[1, 1, 3] == 15.0 and t[1, 1, 17] == 1.0


def test_h09_dummy_neighbors(pkg, orc):
    g = pkg.SparseGrid("h", 32, 16, 16, 0.0, 1.0)
    g.set_voxel(0, 0, 0, 3.0)
    g.set_full_block(1)
    g.set_empty_value(2.0)
    t = orc.halo_fill(g, 0, 1, True)
    assert t[1, 1, 17] == 1.0 and t[1, 1, 0] == 3.0


def test_h10_faces_only_subset_of_full(pkg, orc):
    g = bid_grid(pkg)
    full = orc.halo_fill(g, 13, 1, True)
    faces = orc.halo_fill(g, 13, 1, False)
    idx = np.arange(18)
    xh = (idx == 0) | (idx == 17)
    Z, Y, X = np.meshgrid(idx, idx, idx, indexing="ij")
    hole = xh[X] & (xh[Y] | xh[Z])
    assert np.all(faces[hole] == 0.0)
    assert np.array_equal(full[~hole], faces[~hole])


def test_h12_halo2_fill(pkg, orc):
    g = bid_grid(pkg)
    t = orc.halo_fill(g, 13, 2, True)
    at = lambda x, y, z: t[z, y, x]
    assert at(2, 2, 2) == 13 and at(0, 2, 2) == 12 and at(1, 2, 2) == 12 and at(18, 2, 2) == 14
    assert at(19, 2, 2) == 14 and at(2, 0, 2) == 10 and at(2, 18, 2) == 16 and at(2, 2, 0) == 4
    assert at(2, 2, 18) == 22 and at(0, 0, 0) == 0 and at(19, 19, 19) == 26


def test_h14_multires_coarse_neighbor(pkg, orc):
    fine = pkg.SparseGrid("f", 32, 16, 16, 0.0, 0.0)
    coarse = pkg.SparseGrid("c", 16, 8, 8, 0.0, 0.0, log2_bsx=3)
    fine.ensure_block(0)
    fine.block_data(0)[:] = 1.0
    coarse.ensure_block(1)
    coarse.block_data(1)[:] = 5.0
    levels = np.array([0, 1], np.uint8)
    tile = np.zeros(18 ** 3, np.float32)
    orc.orc_halo_fill_multires(32, 16, 16, 4, fine.slab, fine.b2s, coarse.slab, coarse.b2s, levels, 0, 1, 1, 1,
                               tile)
    t = tile.reshape(18, 18, 18)
    assert t[1, 1, 17] == 5.0 and t[1, 1, 1] == 1.0 and t[1, 1, 0] == 1.0


def test_morton_known_answers(orc):
    assert orc.orc_morton3(0, 0, 0) == 0
    assert orc.orc_morton3(1, 0, 0) == 1 and orc.orc_morton3(0, 1, 0) == 2 and orc.orc_morton3(0, 0, 1) == 4
    assert orc.orc_morton3(1, 1, 1) == 7
    for x, y, z in [(0, 0, 2047), (2047, 0, 0), (1024, 1024, 1024), (2047, 2047, 2047)]:
        c = orc.orc_morton3(x, y, z)
        assert c != 0 and c < (1 << 63)


def test_morton_matches_bit_interleave(pkg, orc):
    """Closed form: bit i of x/y/z lands at bit 3i/3i+1/3i+2 (brute force)."""
    rng = np.random.default_rng(0)
    for x, y, z in rng.integers(0, 1 << 21, size=(500, 3)):
        want = 0
        for i in range(21):
            want |= ((int(x) >> i) & 1) << (3 * i)
            want |= ((int(y) >> i) & 1) << (3 * i + 1)
            want |= ((int(z) >> i) & 1) << (3 * i + 2)
        assert orc.orc_morton3(int(x), int(y), int(z)) == want == pkg.morton3(int(x), int(y), int(z))
