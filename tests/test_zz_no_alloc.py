"""msbg_active_blocks_noalloc against msbg_active_blocks, through the C ABI:
two calls in a row (the second on a smaller input, so stale flags of the
first would show), the largest block grid it accepts, an empty particle set,
and the size limit reported as an error before any device work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run(hip, name, s, pos, r_scan):
    d = hip.to_device(np.ascontiguousarray(pos, np.float32).reshape(-1, 3))
    nb = ((s[0] + 15) >> 4) * ((s[1] + 15) >> 4) * ((s[2] + 15) >> 4)
    out = hip.empty(max(nb, 1), np.int32)
    k = C.c_int64(-1)
    hip.call(name, s[0], s[1], s[2], d.ptr, d.shape[0], r_scan, out.ptr, C.byref(k))
    return out.get()[: k.value].copy()


def both(hip, s, pos, r_scan):
    want = run(hip, "msbg_active_blocks", s, pos, r_scan)
    got = run(hip, "msbg_active_blocks_noalloc", s, pos, r_scan)
    assert np.array_equal(got, want), (got.size, want.size)
    return got


def test_noalloc_matches_allocating_entry_point_twice(hip):
    rng = np.random.default_rng(0)
    s1 = (1024, 1024, 1024)  # 262,144 blocks: the largest grid accepted
    pos1 = rng.uniform(0, 1024, size=(400_000, 3)).astype(np.float32)
    a1 = both(hip, s1, pos1, 3.0)
    assert a1.size > 100_000
    s2 = (200, 96, 130)  # second call: smaller grid, fewer particles
    pos2 = rng.uniform(0, [200, 96, 130], size=(3_000, 3)).astype(np.float32)
    pos2[:10] = [[0, 0, 0]] * 10
    a2 = both(hip, s2, pos2, 2.5)
    assert 0 < a2.size < a1.size and a2.max() < 13 * 6 * 9
    a3 = both(hip, s2, pos2[:1], 4.0)  # one particle: only its own footprint
    assert a3.size == 1


def test_noalloc_empty_and_limit(hip):
    assert both(hip, (64, 64, 64), np.zeros((0, 3), np.float32), 3.0).size == 0
    from msbg_rs_amd.lib import MsbgError
    with pytest.raises(MsbgError) as e:
        run(hip, "msbg_active_blocks_noalloc", (1040, 1024, 1024), np.zeros((1, 3), np.float32), 3.0)
    assert "MSBG_NOALLOC_MAX_BLOCKS" in str(e.value)
