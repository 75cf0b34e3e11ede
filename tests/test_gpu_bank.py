"""GPU parity tests of the bank path: every run's commit log replays bit-exactly.

Each case prints what it runs before touching the device.  The checker is the
CPU oracle (oracle/csmv_oracle.c): serial replay in commit-timestamp order of
every committed transfer, audit snapshots, final balances, conservation.
"""
import numpy as np
import pytest

import csmv
import orc

pytestmark = pytest.mark.gpu

INIT = 1000


def run_bank(n_items, bank, variant="stamp", depth=4, atr=4096, waves=8, wg_waves=8,
             check_stream=True, timeout_ms=4000, route=None, serial_waves=0, begin_wait=True,
             client_wg_waves=0, escalate_after=0, max_attempts=0, commit_parts=0):
    print(f"\n[case] n_items={n_items} bank={bank} variant={variant} depth={depth} atr={atr} "
          f"waves={waves} serial_waves={serial_waves} routed={route is not None} "
          f"client_wg_waves={client_wg_waves} escalate_after={escalate_after} "
          f"max_attempts={max_attempts} commit_parts={commit_parts}", flush=True)
    with csmv.Engine(n_items, depth=depth, atr_entries=atr, client_waves=waves,
                     wg_waves=wg_waves, variant=variant, timeout_ms=timeout_ms,
                     n_serial_waves=serial_waves, begin_wait=begin_wait,
                     client_wg_waves=client_wg_waves, escalate_after=escalate_after,
                     max_attempts=max_attempts, commit_parts=commit_parts) as e:
        e.fill(INIT)
        e.bank_generate(bank)
        if route is not None:
            op, _ = e.bank_read_stream()
            e.bank_route(route(op))
        stats = e.bank_run()
        op, addr = e.bank_read_stream()
        ts, v, ab = e.bank_read_log()
        final = e.read_state()
    spec = orc.bank_spec(n_items, bank.n_txns, seed=bank.seed, n_reads=bank.n_reads,
                         n_reads_ro=bank.n_reads_ro, pct_ro=bank.pct_ro, zipf=bank.zipf,
                         theta=bank.theta, amount_max=bank.amount_max)
    if check_stream:
        oop, oaddr = orc.bank_gen(spec)
        assert np.array_equal(oop, op), "op stream differs from the oracle's"
        assert np.array_equal(oaddr, addr), "operand stream differs from the oracle's"
    rep, _ = orc.bank_replay(spec, op, addr, INIT, ts, v, final)
    print("[stats]", {k: stats[k] for k in ("commits", "aborts", "abort_rate", "kernel_ms",
                                             "txn_per_s", "batches", "escalated")}, flush=True)
    assert rep["fails"] == 0, rep
    assert rep["total"] == n_items * INIT
    assert stats["commits"] == bank.n_txns
    assert stats["commits_update"] == rep["n_updates"]
    assert stats["final_ts"] == rep["n_updates"]
    assert stats["error"] == 0
    assert ab.sum() == stats["aborts"]
    return stats


@pytest.mark.parametrize("variant", list(csmv.VARIANTS))
def test_bank_plain_small(variant):
    """Config 'bank plain' shape, reduced stream: 4096 accounts, 2-account transfers."""
    run_bank(4096, csmv.Bank(n_txns=1 << 15, seed=11), variant=variant, waves=8)


@pytest.mark.parametrize("variant", list(csmv.VARIANTS))
def test_bank_zipf_small(variant):
    """Config 'bank at scale' shape at 2^16 accounts: Zipf 0.99, 4 reads / 2 writes."""
    run_bank(1 << 16, csmv.Bank(n_txns=1 << 15, seed=5, n_reads=4, n_reads_ro=4, zipf=True),
             variant=variant, depth=8, waves=16)


BANK_AT_SCALE = dict(n_reads=4, n_reads_ro=4, pct_ro=10, zipf=True, theta=0.99)


@pytest.mark.parametrize("variant", ["vanilla", "batchedAtrInsert", "clientWriteback",
                                     "collabValidation"])
def test_bank_at_scale_reference_variants(variant):
    """The bank-at-scale configuration's four reference variants at its shape: 2^20
    accounts, Zipf 0.99, transfers reading 4 / writing 2, 4-read audits, 90/10, depth 8,
    record ring 4096, 64 client waves in arrival order, on a 2^14-transaction stream (long
    enough that most attempts abort)."""
    st = run_bank(1 << 20, csmv.Bank(n_txns=1 << 14, seed=13, **BANK_AT_SCALE),
                  variant=variant, depth=8, atr=4096, waves=64, wg_waves=5, client_wg_waves=1,
                  timeout_ms=20000)
    assert st["aborts"] > 0
    assert st["escalated"] == 0


def test_bank_at_scale_stamp_escalation():
    """The stamp variant's normal lane at the bank-at-scale shape with forward progress:
    a transaction that has failed two attempts is executed by the commit server on the
    serialized lane.  The log replays and the escalated transactions are counted."""
    st = run_bank(1 << 20, csmv.Bank(n_txns=1 << 16, seed=14, **BANK_AT_SCALE),
                  variant="stamp", depth=8, waves=64, wg_waves=5, client_wg_waves=1,
                  escalate_after=2, timeout_ms=20000)
    assert st["escalated"] > 0


def test_escalation_instead_of_livelock():
    """A hot stream with an attempt bound of one: without escalation the run stops with
    the attempt-limit error; with escalation after one failed attempt every transaction
    commits and no transaction fails more than once."""
    bank = csmv.Bank(n_txns=1 << 14, seed=15, n_reads=4, n_reads_ro=4, pct_ro=10, zipf=True,
                     theta=1.2)
    print("\n[case] arrival order, max_attempts=1 (expect the attempt-limit error)", flush=True)
    with csmv.Engine(4096, depth=4, client_waves=16, wg_waves=6, variant="stamp",
                     max_attempts=1) as e:
        e.fill(INIT)
        e.bank_generate(bank)
        with pytest.raises(csmv.CsmvError) as ei:
            e.bank_run()
        assert ei.value.status == 7
    st = run_bank(4096, bank, variant="stamp", depth=4, waves=16, wg_waves=6,
                  escalate_after=1, max_attempts=1)
    assert st["escalated"] > 0


@pytest.mark.parametrize("variant", ["batchedAtrInsert", "collabValidation"])
def test_record_age_aborts(variant):
    """A record ring much smaller than the in-flight window forces RecordAge aborts
    (bank/csmv_sh/API.cu:93-100), also while receivers scan records the committer is
    overwriting (collaborative validation at staging); the log must still replay."""
    st = run_bank(4096, csmv.Bank(n_txns=1 << 14, seed=3), variant=variant,
                  atr=64, waves=32)
    assert st["aborts_recordage"] > 0


def test_data_age_aborts():
    """Depth 1 keeps only the newest version, so stale snapshots abort (DataAge)."""
    st = run_bank(256, csmv.Bank(n_txns=1 << 14, seed=4, n_reads_ro=16, pct_ro=50),
                  variant="stamp", depth=1, waves=16)
    assert st["aborts_dataage"] > 0


def test_single_wave_all_commit_order():
    """One client wave: batches arrive one at a time; every transfer commits exactly once."""
    run_bank(1 << 16, csmv.Bank(n_txns=1 << 12, seed=9), variant="stamp", waves=1)


def test_bank_plain_full():
    """Config 'bank plain' at full size: 2^20 transactions, 64 client waves, depth 4."""
    run_bank(4096, csmv.Bank(n_txns=1 << 20, seed=1), variant="stamp", waves=64,
             check_stream=False)


def test_four_wave_server_workgroup():
    """Two receivers instead of six: the staging pipeline depth changes, results may not."""
    run_bank(4096, csmv.Bank(n_txns=1 << 15, seed=21), variant="stamp", waves=16, wg_waves=4)


def test_hashed_stamps_zipf():
    """More items than LDS stamp buckets: hashed stamps with exact HBM stamps behind."""
    run_bank(1 << 20, csmv.Bank(n_txns=1 << 16, seed=8, n_reads=4, n_reads_ro=4, zipf=True),
             variant="stamp", depth=8, waves=64)


def test_no_begin_wait():
    """Snapshots that may predate the wave's own last commit (the collabValidation
    TXBegin, which does not wait) still replay."""
    run_bank(4096, csmv.Bank(n_txns=1 << 15, seed=31), variant="stamp", waves=16,
             begin_wait=False)


def _random_route(frac, seed):
    def f(op):
        return (np.random.default_rng(seed).random(op.size) < frac).astype(np.uint8)
    return f


@pytest.mark.parametrize("frac", [0.5, 1.0])
def test_serialized_lane_exact(frac):
    """Transfers routed to the serialized lane run on the commit server, interleaved
    with normal-lane commits in one timestamp order (exact LDS value table)."""
    st = run_bank(4096, csmv.Bank(n_txns=1 << 16, seed=41), variant="stamp", waves=16,
                  serial_waves=4, route=_random_route(frac, 1))
    assert st["commits_update"] > 0


def test_serialized_lane_hot_zipf_hashed():
    """1M accounts (hashed stamps, HBM value table): transfers touching the 256
    hottest accounts go to the serialized lane."""
    def route(op):
        return None
    n = 1 << 20
    bank = csmv.Bank(n_txns=1 << 17, seed=43, n_reads=4, n_reads_ro=4, zipf=True)
    hot_items = np.array([orc.scramble(r, n) for r in range(256)], dtype=np.uint32)

    def hot_route(op):
        T = op.size
        spec = orc.bank_spec(n, T, seed=bank.seed, n_reads=4, n_reads_ro=4, zipf=True)
        _, addr = orc.bank_gen(spec)
        A = addr.reshape(-1, T)
        return np.isin(A, hot_items).any(axis=0).astype(np.uint8)
    run_bank(n, bank, variant="stamp", depth=8, waves=32, serial_waves=8, route=hot_route,
             wg_waves=6, check_stream=False)


@pytest.mark.parametrize("n_items,pct_ro", [(4096, 30), (1 << 20, 30), (1 << 16, 100)])
def test_serialized_lane_read_only(n_items, pct_ro):
    """Every transaction, audits included, on the serialized lane: an audit is read in
    commit order without a timestamp of its own and logged with the snapshot it saw (the
    last commit below it); nothing aborts, and the commit timestamps stay dense."""
    bank = csmv.Bank(n_txns=1 << 16, seed=53, n_reads=4, n_reads_ro=4, pct_ro=pct_ro,
                     zipf=True)
    st = run_bank(n_items, bank, variant="stamp", depth=8, waves=16, serial_waves=8,
                  wg_waves=6, route=lambda op: np.ones(op.size, dtype=np.uint8))
    assert st["aborts"] == 0
    assert st["commits_ro"] > 0
    if pct_ro == 100:
        assert st["commits_update"] == 0


def test_serialized_lane_small_cache_unknowns():
    """A tiny LDS budget leaves few stamp buckets, so more operands are flagged and resolved
    from the HBM value table: results must not change."""
    n = 1 << 16
    bank = csmv.Bank(n_txns=1 << 15, seed=47, n_reads=4, n_reads_ro=4, zipf=True)
    print("\n[case] lds_kib=64 serialized lane", flush=True)
    with csmv.Engine(n, depth=8, client_waves=16, wg_waves=4, variant="stamp",
                     n_serial_waves=4, lds_kib=64) as e:
        e.fill(INIT)
        e.bank_generate(bank)
        op, addr = e.bank_read_stream()
        e.bank_route(((op & 0xFF) == 0).astype(np.uint8))
        st = e.bank_run()
        ts, v, ab = e.bank_read_log()
        final = e.read_state()
    spec = orc.bank_spec(n, bank.n_txns, seed=bank.seed, n_reads=4, n_reads_ro=4, zipf=True)
    rep, _ = orc.bank_replay(spec, op, addr, INIT, ts, v, final)
    print("[stats]", {k: st[k] for k in ("commits", "aborts_unknown", "stamp_exact")}, flush=True)
    assert rep["fails"] == 0, rep
    assert st["stamp_exact"] == 0


@pytest.mark.parametrize("client_wg_waves,route_all", [(1, False), (2, True)])
def test_clients_spread_over_workgroups(client_wg_waves, route_all):
    """Client waves spread over more workgroups (one or two per workgroup, the other waves
    of each leave at once), at the bank-at-scale shape with the transfers, or everything, on
    the serialized lane: the log replays and every transaction commits."""
    bank = csmv.Bank(n_txns=1 << 17, seed=59, n_reads=4, n_reads_ro=4, pct_ro=10, zipf=True)
    route = ((lambda op: np.ones(op.size, dtype=np.uint8)) if route_all
             else (lambda op: ((op & 0xFF) == 0).astype(np.uint8)))
    st = run_bank(1 << 20, bank, variant="stamp", depth=8, waves=24, serial_waves=6,
                  wg_waves=5, route=route, client_wg_waves=client_wg_waves)
    if route_all:
        assert st["aborts"] == 0


def test_empty_stream_is_rejected():
    """A stream of no transactions is an argument error, raised before any launch."""
    print("\n[case] empty bank stream", flush=True)
    with csmv.Engine(4096, depth=4, client_waves=8) as e:
        with pytest.raises(csmv.CsmvError):
            e.bank_generate(csmv.Bank(n_txns=0, seed=1))


@pytest.mark.parametrize("n_txns,routed", [(1, False), (1, True), (1000, False), (1000, True),
                                           (4097, True)])
def test_ragged_streams(n_txns, routed):
    """Streams that do not fill the last wave (one transaction, 1000, 4097), on the normal
    lane and with every transfer on the serialized lane: the log replays exactly."""
    bank = csmv.Bank(n_txns=n_txns, seed=67, n_reads=4, n_reads_ro=4, pct_ro=10, zipf=True)
    route = (lambda op: ((op & 0xFF) == 0).astype(np.uint8)) if routed else None
    run_bank(1 << 16, bank, variant="stamp", depth=8, waves=12, wg_waves=5,
             serial_waves=4 if routed else 0, route=route, client_wg_waves=1)


def test_run_logs_to_files_replay(tmp_path):
    """A run's commit log and abort log written to files (csmv_log_write_bank /
    csmv_log_write_aborts) and read back replay exactly and keep every count."""
    bank = csmv.Bank(n_txns=1 << 15, seed=91, n_reads=4, n_reads_ro=4, zipf=True)
    print("\n[case] logs through files", flush=True)
    with csmv.Engine(1 << 16, depth=8, client_waves=16, wg_waves=6, variant="stamp",
                     escalate_after=2) as e:
        e.fill(INIT)
        e.bank_generate(bank)
        st = e.bank_run()
        op, addr = e.bank_read_stream()
        ts, v, ab = e.bank_read_log()
        final = e.read_state()
        e.abort_fold(32)
        counters = e.abort_counters()
    pc, pa = str(tmp_path / "commits.log"), str(tmp_path / "aborts.log")
    csmv.log_write_bank(pc, ts, v, ab, seed=bank.seed)
    csmv.log_write_aborts(pa, st, ab, counters, seed=bank.seed)
    h, ts2, v2, ab2 = csmv.log_read_bank(pc)
    _, classes, att, cnt = csmv.log_read_aborts(pa)
    spec = orc.bank_spec(1 << 16, bank.n_txns, seed=h.seed, n_reads=h.rows, n_reads_ro=4,
                         zipf=True)
    rep, _ = orc.bank_replay(spec, op, addr, INIT, ts2, v2, final)
    assert rep["fails"] == 0, rep
    assert np.array_equal(ab2, ab) and np.array_equal(att, ab)
    assert np.array_equal(cnt, counters)
    assert classes["commits_update"] + classes["commits_ro"] == bank.n_txns
    assert classes["escalated"] == st["escalated"]
    assert int(att.sum()) == st["aborts"]


# ---- pipelined serialized lane (variant stamp_pipe) -----------------------------------------
def _transfers(op):
    return ((op & 0xFF) == 0).astype(np.uint8)


def _everything(op):
    return np.ones(op.size, dtype=np.uint8)


@pytest.mark.parametrize("n_items,depth,route,parts,wg",
                         [(4096, 4, "transfers", 1, 6), (4096, 4, "transfers", 4, 8),
                          (1 << 20, 8, "transfers", 1, 8), (1 << 20, 8, "transfers", 4, 8),
                          (1 << 20, 8, "all", 2, 12), (1 << 16, 8, "all", 3, 5),
                          (1 << 20, 8, "transfers", 4, 16), (1 << 20, 8, "all", 8, 16),
                          (1 << 16, 8, "transfers", 6, 8), (4096, 1, "all", 8, 16),
                          (1 << 20, 64, "transfers", 7, 16), (1 << 16, 2, "all", 5, 12),
                          (1 << 20, 1, "transfers", 8, 16)])
def test_pipelined_lane_replays(n_items, depth, route, parts, wg):
    """The pipelined serialized lane: clients stage their requests' chains, committer waves
    apply each item's request-wide delta with one ordered atomic and install the versions
    themselves, split over commit partitions by item hash.  Every transfer (or every
    transaction) on it, at the bank-plain shape and the bank-at-scale shape (1M Zipf 0.99
    accounts, 4 reads / 2 writes, depth 8), and at depth 1 (every install of an account
    writes the same slot, so installs landing out of request order would leave an older
    version there) and depth 64: the log replays exactly and the audits on the normal lane
    read consistent snapshots."""
    zipf = n_items > 4096
    bank = csmv.Bank(n_txns=1 << 17, seed=71, n_reads=4 if zipf else 2,
                     n_reads_ro=4 if zipf else 32, pct_ro=10, zipf=zipf)
    st = run_bank(n_items, bank, variant="stamp_pipe", depth=depth, waves=24, serial_waves=16,
                  wg_waves=wg, client_wg_waves=1, commit_parts=parts,
                  route=_transfers if route == "transfers" else _everything)
    if route == "all" and bank.n_reads_ro <= 4:
        # (audits of more reads than the lane's wire carries stay on the normal lane)
        assert st["aborts"] == 0


def test_pipelined_lane_at_bench_shape():
    """The benchmark's configuration on a 2^20-transaction stream of its workload: 1M Zipf
    0.99 accounts, 4 reads / 2 writes, 4-read audits, depth 8, every transaction on the
    pipelined lane, 8 commit server workgroups of 16 waves, 248 client waves of which 247 feed
    the lane (four transactions a lane, requests of up to 256): the log replays exactly and
    nothing aborts."""
    bank = csmv.Bank(n_txns=1 << 20, seed=1, n_reads=4, n_reads_ro=4, pct_ro=10, zipf=True)
    st = run_bank(1 << 20, bank, variant="stamp_pipe", depth=8, waves=248, serial_waves=247,
                  wg_waves=16, client_wg_waves=1, commit_parts=8, route=_everything)
    assert st["aborts"] == 0


@pytest.mark.parametrize("n_txns,parts", [(1, 1), (777, 2), (4097, 4), (4097, 8)])
def test_pipelined_lane_ragged(n_txns, parts):
    """Streams that do not fill the last wave, and a single serialized-lane wave."""
    bank = csmv.Bank(n_txns=n_txns, seed=73, n_reads=4, n_reads_ro=4, pct_ro=10, zipf=True)
    run_bank(1 << 16, bank, variant="stamp_pipe", depth=8, waves=4, serial_waves=1, wg_waves=5,
             client_wg_waves=1, route=_transfers, commit_parts=parts)


def test_pipelined_lane_rejects_shared_client_workgroups():
    """A lane-feeding client wave stages four transactions a lane in LDS: two of them do not
    fit one workgroup, and the run is refused with a capacity error before any launch."""
    bank = csmv.Bank(n_txns=1 << 12, seed=83, n_reads=4, n_reads_ro=4, pct_ro=10, zipf=True)
    print("\n[case] pipelined lane with two client waves per workgroup", flush=True)
    with csmv.Engine(1 << 16, depth=8, client_waves=8, wg_waves=8, variant="stamp_pipe",
                     n_serial_waves=6, client_wg_waves=2) as e:
        e.fill(INIT)
        e.bank_generate(bank)
        op, _ = e.bank_read_stream()
        e.bank_route(_everything(op))
        with pytest.raises(csmv.CsmvError) as ei:
            e.bank_run()
        assert ei.value.status == 3


def test_pipelined_lane_needs_every_transfer_routed():
    """A transfer left on the normal lane reaches the pipelined server as a normal-lane
    update: the run ends with an argument error instead of committing it unvalidated."""
    bank = csmv.Bank(n_txns=1 << 12, seed=79, n_reads=4, n_reads_ro=4, pct_ro=10, zipf=True)
    print("\n[case] pipelined lane with half the transfers unrouted", flush=True)
    with csmv.Engine(1 << 16, depth=8, client_waves=8, wg_waves=5, variant="stamp_pipe",
                     n_serial_waves=4, client_wg_waves=1) as e:
        e.fill(INIT)
        e.bank_generate(bank)
        op, _ = e.bank_read_stream()
        half = _transfers(op) & (np.arange(op.size) % 2 == 0).astype(np.uint8)
        e.bank_route(half)
        with pytest.raises(csmv.CsmvError) as ei:
            e.bank_run()
        assert ei.value.status == 2
