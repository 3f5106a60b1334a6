"""GPU tests of the grid payload transfers (staged batch upload / download
through a scatter / gather kernel), the device deep copy and the slot-table
accessor: exact round trips against host data."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("log2_bsx,shape", [(4, (64, 48, 40)), (3, (40, 24, 32))])
def test_write_read_blocks_roundtrip(pkg, hip, log2_bsx, shape):
    from msbg_rs_amd.lib import MsbgError
    rng = np.random.default_rng(log2_bsx)
    g = pkg.DeviceGrid(*shape, log2_bsx, 0.25, 0.75, capacity_blocks=512, lib=hip)
    n = g.n_blocks
    vb = rng.permutation(n)[: n * 2 // 3]
    g.ensure_blocks_parallel(vb)
    full = np.setdiff1d(np.arange(n), vb)[:2]
    for b in full:
        g.set_full_block(int(b))
    part = rng.permutation(vb)[: vb.size // 2]
    data = rng.normal(size=(part.size, g.N)).astype(np.float32)
    g.write_blocks(part, data)
    assert np.array_equal(g.read_blocks(part), data)
    # a second batch in another order leaves the first untouched
    rest = np.setdiff1d(vb, part)
    more = rng.normal(size=(rest.size, g.N)).astype(np.float32)
    g.write_blocks(rest[::-1], more[::-1])
    assert np.array_equal(g.read_blocks(rest), more)
    assert np.array_equal(g.read_blocks(part), data)
    # dummies read their sentinel values; writing one is an error
    absent = np.setdiff1d(np.arange(n), np.concatenate([vb, full]))[:1]
    got = g.read_blocks(np.concatenate([full[:1], absent]))
    assert np.all(got[0] == 0.75) and np.all(got[1] == 0.25)
    with pytest.raises(MsbgError):
        g.write_blocks(full[:1], np.zeros((1, g.N), np.float32))
    # the host copy agrees block by block
    h = g.to_host()
    assert np.array_equal(h.slab[h.b2s[part]], data)


def test_grid_copy_and_block_slots(pkg, hip):
    rng = np.random.default_rng(5)
    g = pkg.DeviceGrid(48, 48, 48, 4, 0.0, 1.0, capacity_blocks=27, lib=hip)
    vb = rng.permutation(27)[:20]
    g.ensure_blocks_parallel(vb)
    g.write_blocks(vb, rng.normal(size=(20, 4096)).astype(np.float32))
    b2s = g.block_slots()
    assert np.array_equal(np.sort(np.nonzero(b2s >= 2)[0]), np.sort(vb))
    assert np.array_equal(g.value_blocks(), np.sort(vb))
    c = pkg.DeviceGrid(48, 48, 48, 4, 0.0, 1.0, capacity_blocks=27, lib=hip)
    c.copy_from(g)
    a, b = g.to_host(), c.to_host()
    assert np.array_equal(a.b2s, b.b2s) and np.array_equal(a.slab, b.slab)
    small = pkg.DeviceGrid(48, 48, 48, 4, 0.0, 1.0, capacity_blocks=4, lib=hip)
    from msbg_rs_amd.lib import MsbgError
    with pytest.raises(MsbgError):
        small.copy_from(g)


def test_fill_u32(pkg, hip):
    """msbg_fill_u32 writes exactly the words asked for; a misaligned pointer
    is refused before any launch."""
    from msbg_rs_amd.lib import MsbgError
    a = hip.to_device(np.arange(1000, dtype=np.uint32))
    hip.call("msbg_fill_u32", a.ptr + 4 * 10, 977, 0xDEADBEEF)
    got = a.get()
    want = np.arange(1000, dtype=np.uint32)
    want[10:987] = 0xDEADBEEF
    assert np.array_equal(got, want)
    hip.call("msbg_fill_u32", a.ptr, 0, 7)  # nothing to do
    with pytest.raises(MsbgError):
        hip.call("msbg_fill_u32", a.ptr + 2, 4, 0)
    assert np.array_equal(a.get(), want)
