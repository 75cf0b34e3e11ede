"""The wave64 lane-exchange primitives of include/csmv_device.hpp (DPP scans and
reductions, ds_bpermute gathers) against numpy, through a test kernel built only against
include/ (tests/driver/lane_check.hip)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(nw, rng):
    kinds = []
    v = rng.integers(0, 2**32, size=(nw, 64), dtype=np.uint64).astype(np.uint32)
    # waves of special values: all equal, all zero, all max, one small value at each end,
    # small values (sums without wrap), and sums that wrap past 2^32
    v[0] = 7
    v[1] = 0
    v[2] = 0xFFFFFFFF
    v[3] = 0xFFFFFFFF
    v[3, 0] = 5
    v[4] = 0xFFFFFFFF
    v[4, 63] = 5
    v[5] = rng.integers(0, 1000, size=64)
    v[6, :] = 0x80000000
    v64 = rng.integers(0, 2**64, size=(nw, 64), dtype=np.uint64)
    v64 &= rng.integers(0, 2**64, size=(nw, 64), dtype=np.uint64)   # sparser bits
    v64[0] = 0
    v64[1] = 0
    v64[1, 17] = np.uint64(1) << np.uint64(63)
    v64[2] = 0
    v64[2, 63] = 1
    for b in range(64):
        v64[3, b] = np.uint64(1) << np.uint64(b)
    src = rng.integers(0, 64, size=(nw, 64), dtype=np.uint32)
    src[0] = np.arange(64)
    src[1] = 63 - np.arange(64)
    src[2] = 0
    return v, v64, src


def test_lane_primitives_match_numpy():
    import c1drv
    nw = 1024
    rng = np.random.default_rng(5)
    v, v64, src = _inputs(nw, rng)
    print(f"\n[case] lane primitives over {nw} waves", flush=True)
    scan, gather, red = c1drv.lane_check(v.ravel(), v64.ravel(), src.ravel())
    scan = scan.reshape(nw, 64)
    gather = gather.reshape(nw, 64)
    want_scan = (np.cumsum(v.astype(np.uint64), axis=1) & 0xFFFFFFFF).astype(np.uint32)
    assert np.array_equal(scan, want_scan)
    assert np.array_equal(gather, np.take_along_axis(v, src.astype(np.int64), axis=1))
    assert np.array_equal(red[:, 0], v.min(axis=1))
    assert np.array_equal(red[:, 1], v.max(axis=1))
    assert np.array_equal(red[:, 2], want_scan[:, 63])
    o = np.bitwise_or.reduce(v64, axis=1)
    assert np.array_equal(red[:, 3], (o & np.uint64(0xFFFFFFFF)).astype(np.uint32))
    assert np.array_equal(red[:, 4], (o >> np.uint64(32)).astype(np.uint32))
