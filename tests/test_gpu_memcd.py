"""GPU parity tests of the memcached-style cache: the request log replays serially in
commit order to the same results and final cache contents (oracle/csmv_oracle.c)."""
import numpy as np
import pytest

import csmv
import orc

pytestmark = pytest.mark.gpu


def run_memcd(m: csmv.Memcd, waves=16, wg_waves=8, depth=4, check_stream=True,
              client_wg_waves=0):
    print(f"\n[case] memcd {m} waves={waves} wg_waves={wg_waves} depth={depth} "
          f"client_wg_waves={client_wg_waves}", flush=True)
    with csmv.Engine(m.n_items, depth=depth, client_waves=waves, wg_waves=wg_waves,
                     variant="stamp", timeout_ms=8000, client_wg_waves=client_wg_waves) as e:
        e.fill(0)
        e.memcd_generate(m)
        st = e.memcd_run()
        key, op = e.memcd_read_stream()
        log = e.memcd_read_log()
        final = e.read_state()
    spec = orc.memcd_spec(m.n_reqs, seed=m.seed, n_keys=m.n_keys, n_sets=m.n_sets,
                          n_ways=m.n_ways, pct_get=m.pct_get, batch=m.batch, zipf=m.zipf,
                          theta=m.theta)
    if check_stream:
        okey, oop = orc.memcd_gen(spec)
        assert np.array_equal(okey, key) and np.array_equal(oop, op), "stream differs"
    rep = orc.memcd_replay(spec, key, op, log["ts"], log["res"], log["vsum"], final)
    print("[stats]", {k: st[k] for k in ("commits", "aborts", "abort_rate", "kernel_ms",
                                         "txn_per_s", "batches")}, rep, flush=True)
    assert rep["fails"] == 0, rep
    assert st["commits"] == m.n_reqs and st["error"] == 0
    return st, rep, log


def test_memcd_small_contended():
    """A small cache (64 sets x 8 ways) and 4096 keys: many evictions and conflicts."""
    st, rep, _ = run_memcd(csmv.Memcd(n_reqs=1 << 15, seed=3, n_keys=4096, n_sets=64,
                                      pct_get=80, batch=2048), waves=8)
    assert rep["sets"] > 0 and rep["hits"] > 0


def test_memcd_uniform_keys():
    run_memcd(csmv.Memcd(n_reqs=1 << 15, seed=5, n_keys=100_000, n_sets=1 << 10, zipf=False,
                         batch=4096), waves=16)


def test_memcd_config_shape():
    """The memcached config shape (2^17 sets x 8 ways, 10M-key Zipf 0.99 keyspace,
    95/5) on a reduced stream of 2^18 requests."""
    st, rep, log = run_memcd(csmv.Memcd(n_reqs=1 << 18, seed=1), waves=32, check_stream=False)
    assert rep["hits"] > 0
    assert (log["lat_ticks"] > 0).all()


def test_memcd_one_client_wave_per_workgroup():
    """The cache driver with each client wave in a workgroup of its own (the layout the
    bench uses): results replay exactly."""
    m = csmv.Memcd(n_reqs=1 << 16, seed=61, n_keys=50_000, n_sets=1 << 11, batch=8192)
    run_memcd(m, waves=6, wg_waves=6, client_wg_waves=1)


@pytest.mark.parametrize("n_reqs", [1, 777])
def test_memcd_ragged_streams(n_reqs):
    """Request streams that do not fill the last wave: results replay exactly."""
    m = csmv.Memcd(n_reqs=n_reqs, seed=71, n_keys=5_000, n_sets=1 << 8, batch=256)
    run_memcd(m, waves=6, wg_waves=6, client_wg_waves=1)
