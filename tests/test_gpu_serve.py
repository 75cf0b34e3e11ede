"""GPU tests of the persistent serving loop: Poisson arrivals of cache requests mixed with
bank transactions of read sets 2..64, continuous batching, admission by in-flight
version bytes, and latency from arrival to commit.

Checks: every partition's log replays serially in commit order (oracle), every request
commits once after its arrival, the budget is never exceeded, and each result is the same
alone and in the mix.  "The same" is exact (tolerance 0) on what does not depend on the
interleaving of a partition's own requests: GET results (hit, way, value checksum) over a
warmed cache with no SETs in flight, the final cache, and the final bank balances
(transfers commute); what does depend on it (the LRU "wrote" bit of a GET, a transfer's
read values) is checked by the replay of each run.
"""
import dataclasses

import numpy as np
import pytest

import csmv
import orc

pytestmark = pytest.mark.gpu

INIT = 1000
N_ACC = 1 << 16


def cache_engine(m: csmv.Memcd, waves=16, client_wg_waves=0):
    return csmv.Engine(m.n_items, depth=4, client_waves=waves, wg_waves=6, variant="stamp",
                       timeout_ms=8000, client_wg_waves=client_wg_waves)


def bank_engine(waves=8, client_wg_waves=0):
    return csmv.Engine(N_ACC, depth=8, client_waves=waves, wg_waves=6, variant="stamp",
                       timeout_ms=8000, client_wg_waves=client_wg_waves)


def serve_bank(n, seed=7):
    return csmv.Bank(n_txns=n, seed=seed, n_reads=64, n_reads_min=2, pct_ro=0, zipf=True,
                     theta=0.99)


def bank_spec(b, n_acc=N_ACC):
    return orc.bank_spec(n_acc, b.n_txns, seed=b.seed, n_reads=b.n_reads,
                         n_reads_ro=b.n_reads_ro, pct_ro=b.pct_ro, zipf=b.zipf, theta=b.theta,
                         amount_max=b.amount_max, n_reads_min=b.n_reads_min)


def memcd_spec(m):
    return orc.memcd_spec(m.n_reqs, seed=m.seed, n_keys=m.n_keys, n_sets=m.n_sets,
                          n_ways=m.n_ways, pct_get=m.pct_get, batch=m.batch, zipf=m.zipf,
                          theta=m.theta)


def check_cache(e, m, init_state=None):
    key, op = e.memcd_read_stream()
    log = e.memcd_read_log()
    final = e.read_state()
    rep = orc.memcd_replay(memcd_spec(m), key, op, log["ts"], log["res"], log["vsum"], final,
                           init=init_state)
    assert rep["fails"] == 0, rep
    return log, final


def check_bank(e, b, n_acc=N_ACC):
    op, addr = e.bank_read_stream()
    ts, v, ab = e.bank_read_log()
    final = e.read_state()
    spec = bank_spec(b, n_acc)
    oop, oaddr = orc.bank_gen(spec)
    assert np.array_equal(op, oop) and np.array_equal(addr, oaddr), "stream differs"
    nr = (op >> 16) & 0xFF
    assert nr.min() >= 2 and nr.max() <= 64 and nr.max() > 32
    rep, _ = orc.bank_replay(spec, op, addr, INIT, ts, v, final)
    assert rep["fails"] == 0, rep
    assert rep["total"] == n_acc * INIT
    return final


def check_latency(e, st, n):
    arr = e.serve_arrivals()
    lat = e.serve_latency()
    assert np.all(np.diff(arr.astype(np.int64)) >= 0), "arrivals are not ordered"
    assert st["commits"] == n
    assert np.all(lat > 0)
    return arr, lat


@pytest.mark.parametrize("cpw", [0, 1])
def test_serve_mix_replays_and_reports_latency(cpw):
    """The serving mix replays exactly, with full client workgroups (cpw 0) and with one
    client wave per workgroup (cpw 1)."""
    m = csmv.Memcd(n_reqs=1 << 15, seed=3, n_keys=20_000, n_sets=1 << 10, batch=4096)
    b = serve_bank(1 << 13)
    print(f"\n[case] serve mix: cache {m.n_reqs} reqs + bank {b.n_txns} txns, 20M req/s, "
          f"client_wg_waves={cpw}", flush=True)
    with cache_engine(m, client_wg_waves=cpw) as ec, bank_engine(client_wg_waves=cpw) as eb:
        ec.fill(0)
        ec.memcd_generate(m)
        eb.fill(INIT)
        eb.bank_generate(b)
        out = csmv.serve(ec, eb, offered_rps=20e6, pct_cache=80, seed=5)
        print("[serve]", {k: out[k] for k in ("kernel_ms", "admit_stalls", "peak_inflight")},
              flush=True)
        assert out["final_inflight"] == 0
        check_cache(ec, m)
        check_bank(eb, b)
        _, lc = check_latency(ec, out["cache"], m.n_reqs)
        _, lb = check_latency(eb, out["bank"], b.n_txns)
        khz = out["wallclock_khz"]
        print(f"[latency] cache p50 {np.percentile(lc, 50) / khz * 1e3:.1f} us p99 "
              f"{np.percentile(lc, 99) / khz * 1e3:.1f} us | bank p50 "
              f"{np.percentile(lb, 50) / khz * 1e3:.1f} us p99 "
              f"{np.percentile(lb, 99) / khz * 1e3:.1f} us", flush=True)


def test_serve_budget_caps_inflight_bytes():
    m = csmv.Memcd(n_reqs=1 << 14, seed=4, n_keys=10_000, n_sets=1 << 9, pct_get=80,
                   batch=2048)
    b = serve_bank(1 << 12, seed=9)
    budget = 4096
    print(f"\n[case] serve with a {budget}-byte in-flight budget", flush=True)
    with cache_engine(m) as ec, bank_engine() as eb:
        ec.fill(0)
        ec.memcd_generate(m)
        eb.fill(INIT)
        eb.bank_generate(b)
        out = csmv.serve(ec, eb, offered_rps=50e6, pct_cache=80, seed=6, pool_budget=budget)
        print("[serve]", {k: out[k] for k in ("kernel_ms", "admit_stalls", "peak_inflight")},
              flush=True)
        assert 0 < out["peak_inflight"] <= budget
        assert out["admit_stalls"] > 0
        assert out["final_inflight"] == 0
        check_cache(ec, m)
        check_bank(eb, b)
        check_latency(ec, out["cache"], m.n_reqs)
        check_latency(eb, out["bank"], b.n_txns)


def test_each_result_same_alone_and_in_mix():
    warm = csmv.Memcd(n_reqs=1 << 14, seed=21, n_keys=8000, n_sets=1 << 9, pct_get=0,
                      batch=4096)
    gets = csmv.Memcd(n_reqs=1 << 15, seed=22, n_keys=8000, n_sets=1 << 9, pct_get=100,
                      batch=4096)
    b = serve_bank(1 << 13, seed=23)
    print("\n[case] alone vs mix: GET-only cache stream on a warmed cache + bank stream",
          flush=True)
    with cache_engine(warm) as ec, bank_engine() as eb:
        ec.fill(0)
        ec.memcd_generate(warm)
        ec.memcd_run()
        state = ec.read_state()
        ec.memcd_generate(gets)
        runs = {}
        arr = None
        for mode in ("alone", "mix"):
            ec.load(state)
            eb.fill(INIT)
            eb.bank_generate(b)
            if mode == "alone":
                csmv.serve(ec, None, offered_rps=20e6, pct_cache=80, seed=8)
                log, fc = check_cache(ec, gets, init_state=state)
                arr = ec.serve_arrivals()
                csmv.serve(None, eb, offered_rps=20e6, pct_cache=80, seed=8)
                fb = check_bank(eb, b)
            else:
                csmv.serve(ec, eb, offered_rps=20e6, pct_cache=80, seed=8)
                log, fc = check_cache(ec, gets, init_state=state)
                fb = check_bank(eb, b)
                # the cache class arrives at the same ticks alone and in the mix
                assert np.array_equal(arr, ec.serve_arrivals())
            runs[mode] = (log, fc, fb)
    (la, fca, fba), (lm, fcm, fbm) = runs["alone"], runs["mix"]
    mask = np.uint32(~2 & 0xFFFFFFFF)
    assert np.array_equal(la["res"] & mask, lm["res"] & mask)
    assert np.array_equal(la["vsum"], lm["vsum"])
    assert (la["res"] & 1).sum() > gets.n_reqs // 4, "too few hits to be a meaningful check"
    assert np.array_equal(fca, fcm)
    assert np.array_equal(fba, fbm)
    # and both equal the serial execution in arrival order
    spec = memcd_spec(gets)
    okey, oop = orc.memcd_gen(spec)
    fo, _, ro, vo = orc.memcd_serial(spec, okey, oop, init=state)
    assert np.array_equal(ro & mask, lm["res"] & mask) and np.array_equal(vo, lm["vsum"])
    assert np.array_equal(fo, fcm)


def test_serve_serialized_lane_for_routed_bank_transactions():
    """Bank transactions routed to the serialized lane: the commit server executes them
    (transfer pair through producer chains, the other reads from their producer or the
    committed value); the log replays and no routed transaction aborts."""
    m = csmv.Memcd(n_reqs=1 << 14, seed=31, n_keys=20_000, n_sets=1 << 10, batch=4096)
    b = serve_bank(1 << 13, seed=32)
    print("\n[case] serve mix with the bank partition on the serialized lane", flush=True)
    res = {}
    for routed in (False, True):
        with cache_engine(m) as ec, csmv.Engine(N_ACC, depth=8, client_waves=12, wg_waves=6,
                                                variant="stamp", timeout_ms=8000,
                                                n_serial_waves=4) as eb:
            ec.fill(0)
            ec.memcd_generate(m)
            eb.fill(INIT)
            eb.bank_generate(b)
            if routed:
                op, _ = eb.bank_read_stream()
                eb.bank_route(np.ones(b.n_txns, dtype=np.uint8))
            out = csmv.serve(ec, eb, offered_rps=20e6, pct_cache=80, seed=5)
            check_cache(ec, m)
            check_bank(eb, b)
            _, lb = check_latency(eb, out["bank"], b.n_txns)
            _, _, ab = eb.bank_read_log()
            if routed:
                assert ab.sum() == 0 and out["bank"]["aborts"] == 0
            res[routed] = (out["kernel_ms"], out["bank"]["abort_rate"], np.percentile(lb, 99))
            print(f"[serve] routed={routed} kernel {out['kernel_ms']:.1f} ms bank abort "
                  f"{out['bank']['abort_rate']:.3f} p99 {np.percentile(lb, 99) / out['wallclock_khz']:.2f} ms",
                  flush=True)
    assert res[True][2] < res[False][2]


@pytest.mark.parametrize("n_acc,parts,n_txns,cpw", [(1 << 16, 1, 1 << 13, 1), (1 << 16, 2, 1 << 13, 2),
                                                 (1 << 20, 4, 1 << 14, 1), (1 << 16, 3, 777, 1),
                                                 (1 << 20, 8, 1 << 14, 1)])
def test_serve_pipelined_bank_lane(n_acc, parts, n_txns, cpw):
    """The serving mix's bank partition on the pipelined serialized lane (its accounts split
    over `parts` commit server workgroups; each request carries the lanes whose nodes fit
    every partition's list): both partitions replay exactly, nothing aborts, every
    transaction commits once after its arrival and the budget is respected."""
    m = csmv.Memcd(n_reqs=1 << 14, seed=41, n_keys=20_000, n_sets=1 << 10, batch=4096)
    b = serve_bank(n_txns, seed=43)
    budget = 8192
    print(f"\n[case] serve mix, pipelined bank lane: {n_acc} accounts, {parts} partitions, "
          f"{n_txns} txns, client_wg_waves={cpw}", flush=True)
    with cache_engine(m) as ec, csmv.Engine(n_acc, depth=8, client_waves=8, wg_waves=6,
                                            variant="stamp_pipe", timeout_ms=8000,
                                            n_serial_waves=4, commit_parts=parts,
                                            client_wg_waves=cpw) as eb:
        ec.fill(0)
        ec.memcd_generate(m)
        eb.fill(INIT)
        eb.bank_generate(b)
        out = csmv.serve(ec, eb, offered_rps=40e6, pct_cache=80, seed=13, pool_budget=budget)
        print("[serve]", {k: out[k] for k in ("kernel_ms", "admit_stalls", "peak_inflight")},
              flush=True)
        assert out["final_inflight"] == 0
        assert 0 < out["peak_inflight"] <= budget
        check_cache(ec, m)
        check_bank(eb, b, n_acc)
        _, _, ab = eb.bank_read_log()
        assert ab.sum() == 0 and out["bank"]["aborts"] == 0
        _, lb = check_latency(eb, out["bank"], b.n_txns)
        khz = out["wallclock_khz"]
        print(f"[latency] bank p50 {np.percentile(lb, 50) / khz * 1e3:.1f} us p99 "
              f"{np.percentile(lb, 99) / khz * 1e3:.1f} us", flush=True)


def test_serve_at_config_shape():
    """The serving configuration's shape: the cache configuration (2^17 sets x 8 ways,
    10M-key Zipf 0.99 keyspace, 95/5 GET/SET) beside a 2^20-account Zipf 0.99 bank
    partition with read sets of 2..64, at 50M req/s offered (80/20), with an in-flight
    version budget and bank transactions escalated to the serialized lane after two failed
    attempts.  Both partitions replay exactly, every request commits once after its
    arrival, the budget is never exceeded and admission engages."""
    n_acc = 1 << 20
    m = csmv.Memcd(n_reqs=1 << 17, seed=1)
    b = csmv.Bank(n_txns=1 << 14, seed=7, n_reads=64, n_reads_min=2, pct_ro=0, zipf=True,
                  theta=0.99)
    budget = 16384
    print(f"\n[case] serving at the config shape: cache {m.n_reqs} reqs ({m.n_sets} sets x "
          f"{m.n_ways} ways), bank {b.n_txns} txns over {n_acc} accounts, 50M req/s offered, "
          f"budget {budget}, escalate_after=2", flush=True)
    with cache_engine(m) as ec, csmv.Engine(n_acc, depth=8, client_waves=8, wg_waves=6,
                                            variant="stamp", timeout_ms=20000,
                                            escalate_after=2) as eb:
        ec.fill(0)
        ec.memcd_generate(m)
        eb.fill(INIT)
        eb.bank_generate(b)
        out = csmv.serve(ec, eb, offered_rps=50e6, pct_cache=80, seed=11, pool_budget=budget)
        print("[serve]", {k: out[k] for k in ("kernel_ms", "admit_stalls", "peak_inflight")},
              "bank escalated", out["bank"]["escalated"], flush=True)
        assert out["final_inflight"] == 0
        assert 0 < out["peak_inflight"] <= budget
        assert out["admit_stalls"] > 0
        check_cache(ec, m)
        check_bank(eb, b, n_acc)
        _, lc = check_latency(ec, out["cache"], m.n_reqs)
        _, lb = check_latency(eb, out["bank"], b.n_txns)
        khz = out["wallclock_khz"]
        print(f"[latency] cache p50 {np.percentile(lc, 50) / khz * 1e3:.1f} us p99 "
              f"{np.percentile(lc, 99) / khz * 1e3:.1f} us | bank p50 "
              f"{np.percentile(lb, 50) / khz * 1e3:.1f} us p99 "
              f"{np.percentile(lb, 99) / khz * 1e3:.1f} us", flush=True)


def test_serve_bank_escalation_bounds_latency():
    """Without routing, bank transactions of the serving mix that fail two attempts go to
    the serialized lane: every transaction commits, the log replays, and the bank p99
    latency stays far below that of pure arrival order on the same arrivals."""
    m = csmv.Memcd(n_reqs=1 << 14, seed=33, n_keys=20_000, n_sets=1 << 10, batch=4096)
    b = serve_bank(1 << 13, seed=34)
    p99 = {}
    for esc in (0, 2):
        print(f"\n[case] serve mix, bank escalate_after={esc}", flush=True)
        with cache_engine(m) as ec, csmv.Engine(N_ACC, depth=8, client_waves=8, wg_waves=6,
                                                variant="stamp", timeout_ms=20000,
                                                escalate_after=esc) as eb:
            ec.fill(0)
            ec.memcd_generate(m)
            eb.fill(INIT)
            eb.bank_generate(b)
            out = csmv.serve(ec, eb, offered_rps=5e6, pct_cache=80, seed=12)
            check_cache(ec, m)
            check_bank(eb, b)
            _, lb = check_latency(eb, out["bank"], b.n_txns)
            p99[esc] = np.percentile(lb, 99) / out["wallclock_khz"] * 1e3
            print(f"[serve] escalate_after={esc}: bank abort {out['bank']['abort_rate']:.3f} "
                  f"escalated {out['bank']['escalated']} p99 {p99[esc]:.1f} us", flush=True)
            if esc:
                assert out["bank"]["escalated"] > 0
    assert p99[2] < p99[0]


def test_serve_cache_partitions_at_config_shape():
    """The cache configuration (2^17 sets x 8 ways, 10M-key Zipf 0.99, 95/5) split over four
    commit servers, each owning a quarter of the sets, beside the 2^20-account bank partition
    at 50M req/s offered with an in-flight budget.  Each partition is an ordinary cache of
    2^15 sets and its log replays exactly, the bank replays, every request commits once
    after its arrival, the budget holds and admission engages."""
    n_parts, n_acc = 4, 1 << 20
    full = csmv.Memcd(n_reqs=1 << 18, seed=1)
    print(f"\n[case] serving, cache split over {n_parts} commit servers, 50M req/s", flush=True)
    with csmv.Engine(full.n_items, depth=4, client_waves=1, wg_waves=6) as gen:
        gen.memcd_generate(full)
        key, op = gen.memcd_read_stream()
    parts = csmv.memcd_partition(key, op, full.n_sets, n_parts)
    local = dataclasses.replace(full, n_sets=full.n_sets // n_parts)
    b = csmv.Bank(n_txns=1 << 14, seed=7, n_reads=64, n_reads_min=2, pct_ro=0, zipf=True,
                  theta=0.99)
    engines = [cache_engine(local, waves=8, client_wg_waves=1) for _ in range(n_parts)]
    try:
        for e, (k, o) in zip(engines, parts):
            e.fill(0)
            e.memcd_load(local, k, o)
        with csmv.Engine(n_acc, depth=8, client_waves=8, wg_waves=6, variant="stamp",
                         timeout_ms=20000, escalate_after=2) as eb:
            eb.fill(INIT)
            eb.bank_generate(b)
            out = csmv.serve(engines, eb, offered_rps=50e6, pct_cache=80, seed=13,
                             pool_budget=65536)
            print("[serve]", {k: out[k] for k in ("kernel_ms", "admit_stalls", "peak_inflight")},
                  flush=True)
            assert out["final_inflight"] == 0
            assert 0 < out["peak_inflight"] <= 65536
            assert out["admit_stalls"] > 0
            assert out["cache"]["commits"] == key.size
            check_bank(eb, b, n_acc)
            check_latency(eb, out["bank"], b.n_txns)
        for e, (k, o) in zip(engines, parts):
            m = dataclasses.replace(local, n_reqs=int(k.size))
            check_cache(e, m)
            arr = e.serve_arrivals()
            lat = e.serve_latency()
            assert np.all(np.diff(arr.astype(np.int64)) >= 0) and np.all(lat > 0)
    finally:
        for e in engines:
            e.close()
