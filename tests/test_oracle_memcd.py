"""Pins the memcached-style cache oracle with hand-worked cases and invariants."""
import numpy as np

import orc

GET, SET = 0, 1


def _spec(n, sets=2, ways=2, batch=100):
    return orc.memcd_spec(n, n_keys=16, n_sets=sets, n_ways=ways, batch=batch, zipf=False)


def _items(it, spec):
    ns = spec.n_sets * spec.n_ways
    return it.reshape(12, ns)


def test_hand_worked_set_get_evict():
    """1 set x 2 ways, request k runs at clock k + 1: SET A, SET B, GET A (touches A),
    SET C evicts the least recently used way (B), GET B misses, GET C hits and touches."""
    spec = _spec(6, sets=1, ways=2, batch=1)
    A, B, Cc = (orc.lib() and [orc._memcd_lib().orc_memcd_key(r) for r in (1, 2, 3)])
    key = np.array([A, B, A, Cc, B, Cc], dtype=np.uint64)
    op = np.array([SET, SET, GET, SET, GET, GET], dtype=np.uint32)
    items, ts, res, vsum = orc.memcd_serial(spec, key, op)
    I = _items(items, spec)
    # way choice: A -> way 0 (first invalid), B -> way 1, C evicts B's way 1
    assert [(r >> 4) & 0xF for r in res[[0, 1, 3]]] == [0, 1, 1]
    assert res[2] & 1 and res[2] & 2          # GET A hits and refreshes its clock
    assert not (res[4] & 1)                   # GET B misses after the eviction
    assert res[5] & 1 and res[5] & 2          # GET C hits; clock 6 > slot clock 4
    # final contents: way 0 = A (clock 3), way 1 = C (clock 6)
    assert list(I[0]) == [1, 1]
    assert (I[1, 0], I[2, 0]) == (np.int32(np.uint32(A & 0xFFFFFFFF)), np.int32(np.uint32(A >> 32)))
    assert list(I[3]) == [3, 6]
    # value words of C are the SET's words for request 3
    L = orc._memcd_lib()
    assert [np.uint32(I[4 + j, 1]) for j in range(8)] == [L.orc_memcd_value_word(int(Cc), 3, j)
                                                          for j in range(8)]
    # GET checksum = sum of value words * (2j + 1)
    want = sum(L.orc_memcd_value_word(int(Cc), 3, j) * (2 * j + 1) for j in range(8)) % 2**32
    assert vsum[5] == want
    # commit timestamps: writers only (SET, SET, GET-refresh, SET, GET-refresh) -> 1..5
    assert list(ts[res & 2 != 0]) == [1, 2, 3, 4, 5]
    assert ts[4] == 4                          # the miss reads the state after 4 commits


def test_replay_accepts_serial_and_rejects_planted():
    spec = orc.memcd_spec(4000, n_keys=64, n_sets=4, n_ways=2, batch=50, pct_get=70, zipf=True)
    key, op = orc.memcd_gen(spec)
    items, ts, res, vsum = orc.memcd_serial(spec, key, op)
    rep = orc.memcd_replay(spec, key, op, ts, res, vsum, items)
    assert rep["fails"] == 0 and rep["sets"] > 0 and rep["hits"] > 0
    res2 = res.copy()
    k = int(np.flatnonzero((op == GET) & (res & 1 == 1))[3])
    res2[k] ^= 1 << 4                                    # wrong way reported
    assert orc.memcd_replay(spec, key, op, ts, res2, vsum, items)["bad_result"] >= 1
    v2 = vsum.copy(); v2[k] += 1                         # wrong value seen
    assert orc.memcd_replay(spec, key, op, ts, res, v2, items)["bad_result"] >= 1
    it2 = items.copy(); it2[5] ^= 1                      # wrong final contents
    assert orc.memcd_replay(spec, key, op, ts, res, vsum, it2)["bad_final"] == 1
    w = np.flatnonzero(res & 2)
    ts2 = ts.copy(); ts2[w[10]], ts2[w[11]] = ts[w[11]], ts[w[10]]
    r = orc.memcd_replay(spec, key, op, ts2, res, vsum, items)
    assert r["fails"] >= 0   # a swap of two independent writers may be harmless
    ts3 = ts.copy(); ts3[w[10]] = ts3[w[11]]
    assert orc.memcd_replay(spec, key, op, ts3, res, vsum, items)["bad_ts"] >= 1


def test_keys_sets_and_stream():
    spec = orc.memcd_spec(20000, n_keys=1000, n_sets=16, n_ways=8, pct_get=95, zipf=False)
    key, op = orc.memcd_gen(spec)
    assert abs((op == GET).mean() - 0.95) < 0.01
    assert len(np.unique(key)) <= 1000
    L = orc._memcd_lib()
    sets = np.array([L.orc_memcd_set_of(int(k), 16) for k in key[:2000]])
    assert sets.min() >= 0 and sets.max() < 16 and len(np.unique(sets)) == 16


def test_serial_from_initial_cache_and_get_only_order_independence():
    """A GET-only stream over a warmed cache: results (hit, way, value checksum) and the
    final cache do not depend on the order (timestamps end at the max clock); the
    replay from the warm state accepts the serial log, the replay from empty rejects it."""
    warm = orc.memcd_spec(4000, seed=11, n_keys=3000, n_sets=256, n_ways=4, pct_get=0,
                          batch=500)
    wkey, wop = orc.memcd_gen(warm)
    state, _, _, _ = orc.memcd_serial(warm, wkey, wop)
    gets = orc.memcd_spec(3000, seed=12, n_keys=3000, n_sets=256, n_ways=4, pct_get=100,
                          batch=500)
    key, op = orc.memcd_gen(gets)
    assert (op == 0).all()
    f1, ts1, res1, vs1 = orc.memcd_serial(gets, key, op, init=state)
    rng = np.random.default_rng(3)
    order = rng.permutation(gets.n_reqs).astype(np.uint64)
    f2, _, res2, vs2 = orc.memcd_serial(gets, key, op, order=order, init=state)
    assert np.array_equal(res1 & ~np.uint32(2), res2 & ~np.uint32(2))
    assert np.array_equal(vs1, vs2)
    assert np.array_equal(f1, f2)
    assert (res1 & 1).sum() > 0
    rep = orc.memcd_replay(gets, key, op, ts1, res1, vs1, f1, init=state)
    assert rep["fails"] == 0, rep
    bad = orc.memcd_replay(gets, key, op, ts1, res1, vs1, f1)
    assert bad["fails"] > 0
