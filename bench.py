#!/usr/bin/env python3
"""Benchmark of the gfx950 CSMV engine on the 'bank at scale' workload.

Workload: 1,048,576 int32 accounts (initial balance 1000), account choice Zipf
theta=0.99, transfers read 4 and write 2 accounts, 90/10 transfer/audit mix,
audits read 4 accounts, 2^24 transactions, version depth 8.  The stream is
generated on the device before timing (resident in HBM).  One step = one pass of
the persistent kernel over the whole stream; the pool is reloaded between steps
outside the timed kernel.  value = committed transactions / kernel seconds
(HIP events on the engine's stream).

Prints ONE JSON line.  One GPU only: --gpus above 1 exits with a message.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "csmv-gpu-stm_amd"))

METRIC = "committed txns/s + abort rate (bank 1M Zipf); memcd req/s + p99 latency"
HBM_PEAK_GBS = 8000.0
N_ACCOUNTS = 1 << 20
INIT = 1000
REF_VARIANTS = ("vanilla", "batchedAtrInsert", "clientWriteback", "collabValidation")


def bytes_per_commit(stats, n_reads=4, n_writes=2, n_reads_ro=4):
    """Algorithmic HBM bytes per committed transaction: a read costs head word 4 +
    (version, value) 8 = 12 B, an installed version 12 B, request + reply 12 B; an
    update is 12R + 12W + 12, an audit 12R_ro."""
    upd = 12 * n_reads + 12 * n_writes + 12
    ro = 12 * n_reads_ro
    return stats["commits_update"] * upd + stats["commits_ro"] * ro


def load_traffic(commits):
    """HBM bytes of one launch: the per-commit traffic of the committed counter profile
    (FETCH_SIZE and WRITE_SIZE passes, scaled by a copy kernel of known size) times the
    launch's commits."""
    p = os.path.join(ROOT, "profiles", "bank_zipf_1M_traffic.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            return json.load(f)["hbm_bytes_per_commit"] * commits
    except Exception:
        return None


def cpu_baseline(op, addr, n_txns, sample):
    """The CPU restatement of the reference protocol (oracle), on the host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import orc

    threads = max(2, min(16, os.cpu_count() or 2))
    spec = orc.bank_spec(N_ACCOUNTS, n_txns, seed=1, n_reads=4, n_reads_ro=4, pct_ro=10,
                         zipf=True, theta=0.99)
    _, _, _, st = orc.bank_cpu_csmv(spec, op, addr, INIT, n_run=sample, depth=8,
                                    atr_entries=4096, n_clients=threads - 1, pin=False)
    commits = st["commits_update"] + st["commits_ro"]
    aborts = st["aborts_dataage"] + st["aborts_readwrite"] + st["aborts_recordage"]
    return {
        "value": commits / st["seconds"],
        "unit": "txn/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {sample} transactions of the same stream; {threads - 1} client "
                  f"threads + 1 commit-server thread; abort rate "
                  f"{aborts / max(1, aborts + commits):.3f}; {st['seconds']:.2f} s",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--variant", default="stamp")
    ap.add_argument("--waves", type=int, default=int(os.environ.get("CSMV_BENCH_WAVES", "64")))
    ap.add_argument("--serial-waves", type=int, default=16)
    ap.add_argument("--wg-waves", type=int, default=6)
    ap.add_argument("--router", default="transfers", choices=("none", "transfers"),
                    help="none: every transaction on the normal (client) lane; transfers: "
                         "every transfer on the serialized lane")
    ap.add_argument("--txns", type=int, default=1 << 24)
    ap.add_argument("--variant-txns", type=int, default=1 << 18,
                    help="stream length for the side-by-side reference variants")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=1 << 22)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              f"(the CSMV commit path has one global clock per engine)", file=sys.stderr)
        sys.exit(2)

    import csmv

    bank = csmv.Bank(n_txns=args.txns, seed=1, n_reads=4, n_reads_ro=4, pct_ro=10, zipf=True,
                     theta=0.99)

    import numpy as np

    def make(variant, b, router, serial):
        e = csmv.Engine(N_ACCOUNTS, depth=8, atr_entries=4096, client_waves=args.waves,
                        wg_waves=args.wg_waves, variant=variant, timeout_ms=8000,
                        n_serial_waves=serial if router != "none" else 0)
        e.bank_generate(b)
        if router == "transfers":
            op, _ = e.bank_read_stream()
            e.bank_route(((op & 0xFF) == 0).astype(np.uint8))
        return e

    e = make(args.variant, bank, args.router, args.serial_waves)
    for _ in range(args.warmup):
        e.fill(INIT)
        e.bank_run()
    runs = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        e.fill(INIT)
        runs.append(e.bank_run())
    wall = time.perf_counter() - t0
    commits = sum(r["commits"] for r in runs)
    aborts = sum(r["aborts"] for r in runs)
    kernel_s = sum(r["kernel_ms"] for r in runs) * 1e-3
    value = commits / kernel_s
    ms_per_step = kernel_s * 1e3 / args.steps
    atomic_rate = e.atomic_rate(waves=4096, iters=256)
    op, addr = (None, None)
    if not args.no_cpu:
        op, addr = e.bank_read_stream()
    e.close()

    alg_bytes = bytes_per_commit(runs[-1])
    achieved = alg_bytes / (runs[-1]["kernel_ms"] * 1e-3) / 1e9
    traffic = load_traffic(runs[-1]["commits"])
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic}
    # the per-lane protocol spends at least two same-address atomics per committed
    # update (ready bit + commit counter); its ceiling is the measured rate / 2
    plain_ceiling = atomic_rate / 2.0
    variants = {}
    if not args.no_variants:
        vbank = csmv.Bank(n_txns=args.variant_txns, seed=1, n_reads=4, n_reads_ro=4, pct_ro=10,
                          zipf=True, theta=0.99)
        for v in REF_VARIANTS + ("stamp",):
            ev = make(v, vbank, "none", 0)
            ev.fill(INIT)
            r = ev.bank_run()
            ev.close()
            variants[v] = {"txn_per_s": r["txn_per_s"], "abort_rate": r["abort_rate"],
                           "kernel_ms": r["kernel_ms"], "txns": args.variant_txns,
                           "lane": "normal (client-executed CSMV)"}
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "txn/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {
            "workload": "bank at scale: 1,048,576 int32 accounts, Zipf 0.99, transfers read 4 "
                        "/ write 2, audits read 4, 90/10, 2^24 txns, version depth 8",
            "txns_per_step": args.txns,
            "variant": args.variant,
            "router": args.router,
            "client_waves": args.waves,
            "serial_lane_waves": args.serial_waves if args.router != "none" else 0,
            "server_wg_waves": args.wg_waves,
            "client_lanes": args.waves * 64,
            "record_ring": 4096,
        },
        "aborts_by_class": {k: sum(r[k] for r in runs) for k in
                            ("aborts_dataage", "aborts_readwrite", "aborts_prevalid",
                             "aborts_recordage", "aborts_unknown")},
        "committer_busy_frac": sum(r["prof"]["commit_busy"] for r in runs) / (kernel_s * 1e9),
        "abort_rate": aborts / (aborts + commits),
        "wall_s_timed_steps": wall,
        "roofline": roofline,
        "atomics": {"same_address_per_s": atomic_rate,
                    "plain_per_lane_ceiling_txn_per_s": plain_ceiling,
                    "value_over_plain_ceiling": value / plain_ceiling},
        "variants": variants,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(op, addr, args.txns, min(args.cpu_sample, args.txns))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
