"""On-disk log formats (csmv.h "log files"): a 64-byte little-endian header followed by the
arrays.  Round trips through the C-ABI, the raw byte layout, and detection of damaged files.
No device is needed.  The logs written here are the oracle's serial executions (test data);
a read-back log must replay exactly like the one written."""
import os
import struct

import numpy as np
import pytest

import csmv
import orc

INIT = 1000


def _bank_log():
    spec = orc.bank_spec(256, 5000, seed=3, n_reads=4, n_reads_ro=8, pct_ro=20, zipf=True)
    op, addr = orc.bank_gen(spec)
    state, ts, v = orc.bank_serial(spec, op, addr, INIT)
    ab = (np.arange(spec.n_txns, dtype=np.uint32) * 7) % 5
    return spec, op, addr, state, ts, v, ab


def test_bank_log_round_trip_replays(tmp_path):
    spec, op, addr, state, ts, v, ab = _bank_log()
    p = str(tmp_path / "bank.log")
    csmv.log_write_bank(p, ts, v, ab, seed=spec.seed)
    h, ts2, v2, ab2 = csmv.log_read_bank(p)
    assert (h.magic, h.version, h.kind, h.n, h.rows, h.seed) == (
        b"CSMVLOG", 1, csmv.LOG_BANK, spec.n_txns, 4, spec.seed)
    assert np.array_equal(ts, ts2) and np.array_equal(v, v2) and np.array_equal(ab, ab2)
    rep, _ = orc.bank_replay(spec, op, addr, INIT, ts2, v2, state)
    assert rep["fails"] == 0, rep


def test_bank_log_byte_layout(tmp_path):
    """The header fields sit at their documented offsets, little-endian, and the arrays
    follow in order (ts, v rows, aborts), so the file can be read without this library."""
    spec, _, _, _, ts, v, ab = _bank_log()
    p = str(tmp_path / "bank.log")
    csmv.log_write_bank(p, ts, v, ab, seed=0xABCDEF)
    raw = open(p, "rb").read()
    assert len(raw) == 64 + 4 * (ts.size + v.size + ab.size)
    magic, ver, kind, n, rows, buckets, seed, csum = struct.unpack_from("<8sIIQIIQQ", raw, 0)
    assert magic == b"CSMVLOG\0" and (ver, kind, n, rows, buckets, seed) == (
        1, 1, ts.size, 4, 0, 0xABCDEF)
    assert raw[48:64] == bytes(16)
    body = raw[64:]
    assert np.array_equal(np.frombuffer(body, "<u4", ts.size, 0), ts)
    assert np.array_equal(np.frombuffer(body, "<i4", v.size, 4 * ts.size), v)
    assert np.array_equal(np.frombuffer(body, "<u4", ab.size, 4 * (ts.size + v.size)), ab)
    # FNV-1a 64 over the arrays, computed here independently
    h = 0xcbf29ce484222325
    for b in body:
        h = ((h ^ b) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    assert csum == h


def test_cache_log_round_trip_replays(tmp_path):
    spec = orc.memcd_spec(3000, seed=5, n_keys=500, n_sets=16, n_ways=4, pct_get=80, batch=256)
    key, op = orc.memcd_gen(spec)
    items, ts, res, vsum = orc.memcd_serial(spec, key, op)
    p = str(tmp_path / "cache.log")
    log = {"ts": ts, "res": res, "vsum": vsum, "aborts": np.zeros_like(ts)}
    csmv.log_write_cache(p, log, seed=spec.seed)
    h, back = csmv.log_read_cache(p)
    assert h.kind == csmv.LOG_CACHE and h.n == spec.n_reqs
    for k in log:
        assert np.array_equal(np.asarray(log[k], dtype=np.uint32), back[k])
    rep = orc.memcd_replay(spec, key, op, back["ts"], back["res"], back["vsum"], items)
    assert rep["fails"] == 0, rep


@pytest.mark.parametrize("with_counters", [True, False])
def test_abort_log_round_trip(tmp_path, with_counters):
    stats = {k: 1000 + 17 * i for i, k in enumerate(csmv.ABORT_CLASSES)}
    att = np.random.default_rng(1).integers(0, 50, 777).astype(np.uint32)
    cnt = np.arange(csmv.ABORT_BUCKETS, dtype=np.uint32) if with_counters else None
    p = str(tmp_path / "aborts.log")
    csmv.log_write_aborts(p, stats, att, cnt, seed=9)
    h, cl, att2, cnt2 = csmv.log_read_aborts(p)
    assert h.kind == csmv.LOG_ABORTS and h.n == 777 and h.seed == 9
    assert cl == stats
    assert np.array_equal(att, att2)
    if with_counters:
        assert h.buckets == csmv.ABORT_BUCKETS and np.array_equal(cnt, cnt2)
    else:
        assert h.buckets == 0 and cnt2 is None


def _damage(path, offset, xor=0x01):
    b = bytearray(open(path, "rb").read())
    b[offset] ^= xor
    open(path, "wb").write(bytes(b))


def test_damaged_files_are_rejected(tmp_path):
    _, _, _, _, ts, v, ab = _bank_log()
    p = str(tmp_path / "bank.log")
    cases = {"payload bit": 64 + 4 * 100, "magic": 0, "version": 8, "kind": 12}
    for what, off in cases.items():
        csmv.log_write_bank(p, ts, v, ab)
        _damage(p, off)
        with pytest.raises(csmv.CsmvError) as ei:
            csmv.log_read_bank(p)
        assert ei.value.status == 8, what
    # truncated, and with a trailing byte
    csmv.log_write_bank(p, ts, v, ab)
    raw = open(p, "rb").read()
    for bad in (raw[:-3], raw + b"\0"):
        open(p, "wb").write(bad)
        with pytest.raises(csmv.CsmvError) as ei:
            csmv.log_read_bank(p)
        assert ei.value.status == 8
    # a bank log is not a cache log
    csmv.log_write_bank(p, ts, v, ab)
    with pytest.raises(csmv.CsmvError) as ei:
        csmv.log_read_cache(p)
    assert ei.value.status == 8
    with pytest.raises(csmv.CsmvError) as ei:
        csmv.log_read_bank(str(tmp_path / "missing.log"))
    assert ei.value.status == 8
