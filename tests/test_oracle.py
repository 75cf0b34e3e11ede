"""Pins the CPU oracle against published values, library routines, closed forms and
hand-worked cases (the reference ships no tests or golden vectors)."""
import ctypes
import math

import numpy as np
import pytest

import orc

NONE = 0xFFFFFFFF


# ---- random streams --------------------------------------------------------------------
def test_splitmix64_published_sequence():
    # SplitMix64 seeded with 0 (Steele, Lea, Flood 2014; Vigna's splitmix64.c): the
    # i-th output equals mix(i * golden) with our single-call form.
    golden = 0x9E3779B97F4A7C15
    expect = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    got = [orc.splitmix64((i * golden) & 0xFFFFFFFFFFFFFFFF) for i in range(3)]
    assert got == expect


def test_rand_r_fnc_matches_glibc_rand_r():
    # RAND_R_FNC (bank/csmv_sh/API.cuh:37-53) is glibc rand_r's three-step LCG run on a
    # 64-bit state; its output only depends on the low 32 bits, so glibc pins it.
    libc = ctypes.CDLL("libc.so.6")
    libc.rand_r.argtypes = [ctypes.POINTER(ctypes.c_uint)]
    libc.rand_r.restype = ctypes.c_int
    rng = np.random.default_rng(0)
    for s in list(rng.integers(0, 2**63, 200)) + [0, 1, 12345, 2**32 - 1, 2**63 + 5]:
        s = int(s)
        r, ns = orc.rand_r_fnc(s)
        s32 = ctypes.c_uint(s & 0xFFFFFFFF)
        assert r == libc.rand_r(ctypes.byref(s32))
        assert ns & 0xFFFFFFFF == s32.value
        assert r < 2**31


def test_zipf_cdf_closed_forms():
    n = 1000
    cdf = orc.zipf_cdf(n, 0.99)
    assert cdf[-1] == 2**31
    assert (np.diff(cdf.astype(np.int64)) >= 0).all()
    h = math.fsum((i + 1) ** -0.99 for i in range(n))
    assert abs(int(cdf[0]) - 2**31 / h) <= 1.0
    assert abs(int(cdf[9]) - 2**31 * math.fsum((i + 1) ** -0.99 for i in range(10)) / h) <= 1.0
    # theta = 0 is the uniform distribution: cdf[i] = floor(2^31 (i+1) / n)
    u = orc.zipf_cdf(10, 0.0)
    assert [int(x) for x in u] == [(2**31 * (i + 1)) // 10 for i in range(10)]


def test_zipf_rank_brute_force():
    cdf = orc.zipf_cdf(5, 1.2)
    rs = list(range(0, 2**31, 2**31 // 997)) + [int(c) for c in cdf[:-1]] + \
        [int(c) - 1 for c in cdf[:-1]] + [2**31 - 1]
    for r in rs:
        assert orc.zipf_rank(cdf, r) == int(np.searchsorted(cdf, r, side="right"))


def test_scramble_is_a_bijection():
    for n in (8, 4096):
        assert sorted(orc.scramble(i, n) for i in range(n)) == list(range(n))
    assert orc.scramble(7, 1000) == 7


def test_zipf_stream_frequencies():
    spec = orc.bank_spec(1 << 12, 1 << 16, n_reads=4, n_reads_ro=4, zipf=True, theta=0.99)
    _, addr = orc.bank_gen(spec)
    a = addr[addr != NONE]
    counts = np.bincount(a, minlength=spec.n_accounts)
    h = math.fsum((i + 1) ** -0.99 for i in range(spec.n_accounts))
    # the hottest account is rank 1 (index scramble(0) = 0)
    p0 = counts[0] / a.size
    assert abs(p0 - 1 / h) < 0.01


# ---- bank stream -------------------------------------------------------------------------
def test_bank_stream_shape():
    spec = orc.bank_spec(4096, 20000, pct_ro=10)
    op, addr = orc.bank_gen(spec)
    kind, amt, nr = op & 0xFF, (op >> 8) & 0xFF, (op >> 16) & 0xFF
    assert abs((kind == 1).mean() - 0.10) < 0.01
    assert set(np.unique(amt[kind == 0])) == set(range(1, 11))
    assert (nr[kind == 0] == 2).all() and (nr[kind == 1] == 32).all()
    A = addr.reshape(32, -1)
    assert (A[:2] < 4096).all()
    assert (A[2:, kind == 0] == NONE).all() and (A[:, kind == 1] < 4096).all()
    # a transaction is a pure function of (seed, k): a shorter stream is a prefix
    op2, addr2 = orc.bank_gen(orc.bank_spec(4096, 100, pct_ro=10))
    assert (op2 == op[:100]).all() and (addr2.reshape(32, -1) == A[:, :100]).all()


def _hand_stream():
    """3 accounts, init 1000: T0 transfer 0->1 amt 5, T1 transfer 1->1 amt 3 (read own
    write), T2 audit of {0,1,2}, T3 transfer 2->0 amt 10."""
    spec = orc.bank_spec(3, 4, n_reads=2, n_reads_ro=3, pct_ro=0)
    op = np.array([0 | 5 << 8 | 2 << 16, 0 | 3 << 8 | 2 << 16, 1 | 3 << 16,
                   0 | 10 << 8 | 2 << 16], dtype=np.uint32)
    addr = np.array([[0, 1, 0, 2], [1, 1, 1, 0], [NONE, NONE, 2, NONE]],
                    dtype=np.uint32).reshape(-1)
    return spec, op, addr


def test_serial_hand_worked():
    spec, op, addr = _hand_stream()
    state, ts, v = orc.bank_serial(spec, op, addr, 1000)
    assert state.tolist() == [1005, 1005, 990]
    assert ts.tolist() == [1, 2, 2, 3]          # the audit sees the state after 2 commits
    V = v.reshape(2, -1)
    assert V[:, 0].tolist() == [1000, 1000]
    assert V[:, 1].tolist() == [1005, 1002]     # second read of account 1 is its own write
    assert V[0, 2] == 3000
    assert V[:, 3].tolist() == [1000, 995]
    rep, replayed = orc.bank_replay(spec, op, addr, 1000, ts, v, state)
    assert rep["fails"] == 0 and replayed.tolist() == [1005, 1005, 990]


def _valid_log(seed=3, n=64, T=3000):
    spec = orc.bank_spec(n, T, seed=seed, n_reads=2, n_reads_ro=8, pct_ro=20)
    op, addr = orc.bank_gen(spec)
    state, ts, v = orc.bank_serial(spec, op, addr, 1000)
    return spec, op, addr, state, ts, v


def test_replay_accepts_valid_and_rejects_planted_mistakes():
    spec, op, addr, state, ts, v = _valid_log()
    rep, _ = orc.bank_replay(spec, op, addr, 1000, ts, v, state)
    assert rep["fails"] == 0
    upd = np.flatnonzero((op & 0xFF) == 0)
    aud = np.flatnonzero((op & 0xFF) == 1)
    T = spec.n_txns

    v2 = v.copy(); v2[T + upd[10]] += 1                      # a wrong logged read
    assert orc.bank_replay(spec, op, addr, 1000, ts, v2, state)[0]["bad_reads"] >= 1

    ts2 = ts.copy(); ts2[upd[5]] = ts2[upd[6]]               # duplicated commit ts
    assert orc.bank_replay(spec, op, addr, 1000, ts2, v, state)[0]["bad_ts"] >= 1

    # swapping the commit order of two transfers that touch a common account
    A = addr.reshape(-1, T)
    i, j = next((a, b) for a in upd for b in upd if a < b and
                set(A[:2, a]) & set(A[:2, b]))
    ts3 = ts.copy(); ts3[i], ts3[j] = ts[j], ts[i]
    assert orc.bank_replay(spec, op, addr, 1000, ts3, v, state)[0]["fails"] >= 1

    st2 = state.copy(); st2[7] += 1                          # wrong final state
    r = orc.bank_replay(spec, op, addr, 1000, ts, v, st2)[0]
    assert r["bad_final"] == 1

    # an audit whose snapshot is moved to a point where its sum differs
    for k in aud:
        ts4 = ts.copy()
        ts4[k] = ts[k] + 1 if ts[k] < len(upd) else ts[k] - 1
        r = orc.bank_replay(spec, op, addr, 1000, ts4, v, state)[0]
        s_orig = v[k]
        if r["bad_audits"]:
            break
    else:
        pytest.fail("no audit snapshot shift was detected")


def test_conservation_under_any_serial_order():
    spec, op, addr, state, ts, v = _valid_log(seed=9)
    perm = np.random.default_rng(1).permutation(spec.n_txns).astype(np.uint64)
    st2, ts2, v2 = orc.bank_serial(spec, op, addr, 1000, order=perm)
    assert st2.astype(np.int64).sum() == 64 * 1000
    assert orc.bank_replay(spec, op, addr, 1000, ts2, v2, st2)[0]["fails"] == 0


# ---- CPU CSMV restatement -----------------------------------------------------------------
def test_cpu_csmv_one_client_commits_in_stream_order():
    spec, op, addr, state, ts, v = _valid_log(seed=5)
    st, ts2, v2, stats = orc.bank_cpu_csmv(spec, op, addr, 1000, n_clients=1)
    assert stats["aborts_readwrite"] == 0 and stats["aborts_dataage"] == 0
    assert (ts2 == ts).all() and (v2 == v).all() and (st == state).all()


@pytest.mark.parametrize("clients,depth,atr", [(4, 4, 4096), (7, 1, 4096), (7, 4, 64)])
def test_cpu_csmv_replays(clients, depth, atr):
    spec = orc.bank_spec(16, 20000, seed=clients, n_reads=2, n_reads_ro=4, pct_ro=20)
    op, addr = orc.bank_gen(spec)
    st, ts, v, stats = orc.bank_cpu_csmv(spec, op, addr, 1000, n_clients=clients, depth=depth,
                                         atr_entries=atr)
    rep, _ = orc.bank_replay(spec, op, addr, 1000, ts, v, st)
    assert rep["fails"] == 0, rep
    assert stats["commits_update"] + stats["commits_ro"] == spec.n_txns
