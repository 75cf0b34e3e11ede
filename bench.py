#!/usr/bin/env python3
"""Benchmark of the gfx950 CSMV engine.

Headline ("bank at scale"): 1,048,576 int32 accounts (initial balance 1000), account
choice Zipf theta=0.99, transfers read 4 and write 2 accounts, 90/10 transfer/audit mix,
audits read 4 accounts, 2^24 transactions, version depth 8.  The stream is generated on
the device before timing (resident in HBM).  One step = one pass of the persistent kernel
over the whole stream; the pool is reloaded between steps outside the timed kernel.
value = committed transactions / kernel seconds (HIP events on the engine's stream).
Commit path of the headline: the pipelined serialized lane (variant stamp_pipe) with every
transaction routed to it, 4 commit server workgroups of 16 waves (11 committers, a
publisher, 4 receivers each) over a hash split of the accounts, 80 client waves of which
64 feed the lane.

Beside it, in the same JSON line:
  audits_on_normal_lane / one_committer_serialized_lane
             the same stream with audits as snapshot reads on the normal lane, and on
             the serialized lane of one committer wave (clients install their versions)
  variants   the reference's four commit-path variants on the normal (client-executed)
             lane, same workload on a shorter stream
  memcd      the memcached-style cache at its configured shape (2^17 sets x 8 ways,
             10M-key Zipf 0.99 keyspace, 95/5 GET/SET, batches of 65,536): req/s,
             abort rate, pick-up-to-commit latency
  scheduler  the learned contention predictor on the 'bank plain' shape (4096 accounts,
             2-account transfers) at Zipf 0.6 / 0.9 / 1.2: arrival-order CSMV vs the
             learned routing vs routing every transfer
  serving    the persistent serving loop: Poisson arrivals of cache requests (80 %) and
             bank transactions of 2..64 reads over 1M Zipf 0.99 accounts (20 %) at fixed
             offered loads, p50/p99 latency from arrival to commit: the bank class on one
             commit server with and without the predictor, and on the pipelined lane
  roofline / atomics / cpu_baseline   for the headline kernel

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


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def bytes_per_commit(stats, n_reads=4, n_writes=2, n_reads_ro=4):
    """Algorithmic HBM bytes per committed transaction: a read costs head word 4 +
    (version, value) 8 = 12 B, an installed version 12 B, request + reply 12 B; an
    update is 12R + 12W + 12, an audit 12R_ro."""
    upd = 12 * n_reads + 12 * n_writes + 12
    ro = 12 * n_reads_ro
    return stats["commits_update"] * upd + stats["commits_ro"] * ro


def load_traffic(commits, kernel):
    """HBM bytes of one launch: the per-commit traffic of the committed counter profile
    (FETCH_SIZE and WRITE_SIZE passes, scaled by a copy kernel of known size) times the
    launch's commits; None when that profile was taken of another kernel."""
    p = os.path.join(ROOT, "profiles", "bank_zipf_1M_traffic.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("kernel") != kernel:
        return None
    return t["hbm_bytes_per_commit"] * commits


def pct_us(lat_ticks, khz):
    import numpy as np
    if lat_ticks.size == 0:
        return {"p50_us": None, "p99_us": None}
    return {"p50_us": float(np.percentile(lat_ticks, 50)) / khz * 1e3,
            "p99_us": float(np.percentile(lat_ticks, 99)) / khz * 1e3}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(op, addr, n_txns, sample, sample_1t):
    """The CPU restatement of the reference protocol (oracle) on the host cores: nproc - 1
    pinned client threads + 1 commit-server thread on a bounded prefix of the same stream,
    the same with one client thread, and the single-thread serial replay of the whole
    stream in commit order."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc

    cores = len(os.sched_getaffinity(0))
    threads = max(2, cores)
    spec = orc.bank_spec(N_ACCOUNTS, n_txns, seed=1, n_reads=4, n_reads_ro=4, pct_ro=10,
                         zipf=True, theta=0.99)

    def run(n_run, n_clients):
        _, _, _, st = orc.bank_cpu_csmv(spec, op, addr, INIT, n_run=n_run, depth=8,
                                        atr_entries=4096, n_clients=n_clients, pin=True)
        commits = st["commits_update"] + st["commits_ro"]
        aborts = st["aborts_dataage"] + st["aborts_readwrite"] + st["aborts_recordage"]
        return {"txn_per_s": commits / st["seconds"], "seconds": st["seconds"],
                "abort_rate": aborts / max(1, aborts + commits), "txns": n_run,
                "client_threads": n_clients}

    multi = run(sample, threads - 1)
    single = run(sample_1t, 1)
    t0 = time.perf_counter()
    orc.bank_serial(spec, op, addr, INIT)
    serial_s = time.perf_counter() - t0
    return {
        "value": multi["txn_per_s"],
        "unit": "txn/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {sample} transactions of the same stream; {threads - 1} client "
                  f"threads + 1 commit-server thread pinned one per core; abort rate "
                  f"{multi['abort_rate']:.3f}; {multi['seconds']:.2f} s",
        "cpu_model": cpu_model(),
        "nproc": cores,
        "single_client": single,
        "serial_replay": {"txns": n_txns, "seconds": serial_s, "txn_per_s": n_txns / serial_s,
                          "threads": 1},
    }


def cache_bytes(op, res):
    """Algorithmic HBM bytes of committed cache requests (item-granular, 12 B per item read
    or installed: head word 4 + (version, value) 8): every request reads the state, key
    low/high and timestamp of the set's 8 ways (32 items); a GET hit reads the 8 value
    words; a GET that moves the LRU timestamp installs 1 item, a SET 12; a request that
    writes adds 12 B of request and reply."""
    import numpy as np
    get = op == 0
    hit = (res & 1) != 0
    wrote = (res & 2) != 0
    items_read = 32 + 8 * (get & hit)
    items_written = np.where(get, wrote.astype(np.int64), 12 * wrote.astype(np.int64))
    return int((12 * items_read + 12 * items_written + 12 * wrote).sum())


def roofline(nbytes, ms):
    achieved = nbytes / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes": nbytes}


def bank_plain_section(csmv, args):
    """The bank-plain configuration: 4096 uniform accounts of 1000, 2-account transfers and
    32-account audits 90/10, 2^20 transactions, 64 client waves x 64 lanes + the server
    workgroup, depth 4, on the normal lane in arrival order (vanilla = the reference's
    protocol with the serial scan, collabValidation = the same with the receivers' collaborative
    scan) and the stamp variant (with and without escalation)."""
    bank = csmv.Bank(n_txns=args.plain_txns, seed=1, n_reads=2, n_reads_ro=32, pct_ro=10)
    out = {"workload": f"4096 uniform accounts, 2-account transfers, 32-account audits, 90/10, "
                       f"{args.plain_txns} txns, 64 client waves, depth 4"}
    for v, esc in (("vanilla", 0), ("collabValidation", 0), ("stamp", 0), ("stamp", args.escalate_after)):
        name = v if not esc else "stamp_escalate"
        log(f"bank plain {name}")
        with csmv.Engine(4096, depth=4, atr_entries=4096, client_waves=64, wg_waves=6, variant=v,
                         timeout_ms=20000, escalate_after=esc) as e:
            e.fill(INIT)
            e.bank_generate(bank)
            r = e.bank_run()
        out[name] = {"txn_per_s": r["txn_per_s"], "abort_rate": r["abort_rate"],
                     "kernel_ms": r["kernel_ms"], "escalated": r["escalated"]}
    return out


def memcd_section(csmv, args):
    m = csmv.Memcd(n_reqs=args.memcd_reqs, seed=1)
    log(f"memcd: {m.n_reqs} requests, {m.n_sets} sets x {m.n_ways} ways")
    with csmv.Engine(m.n_items, depth=4, client_waves=args.memcd_waves, wg_waves=6,
                     variant="stamp", timeout_ms=20000,
                     client_wg_waves=args.client_wg_waves) as e:
        e.fill(0)
        e.memcd_generate(m)
        st = e.memcd_run()
        lg = e.memcd_read_log()
        _, mop = e.memcd_read_stream()
        khz = e.wallclock_khz()
    return {"req_per_s": st["txn_per_s"], "abort_rate": st["abort_rate"],
            "kernel_ms": st["kernel_ms"], "requests": m.n_reqs, "client_waves": args.memcd_waves,
            "latency_pickup_to_commit": pct_us(lg["lat_ticks"], khz),
            "roofline": roofline(cache_bytes(mop, lg["res"]), st["kernel_ms"]),
            "workload": "2^17 sets x 8 ways, 64-bit keys Zipf 0.99 over 10M, 32-byte values, "
                        "95/5 GET/SET, LRU clock per batch of 65,536"}


def scheduler_section(csmv, args):
    """C4: train on abort logs at theta 0.6 / 0.9 / 1.2, calibrate, evaluate."""
    import numpy as np
    from csmv import sched

    def eng(serial):
        return csmv.Engine(4096, depth=4, atr_entries=4096, client_waves=64, wg_waves=8,
                           variant="stamp", timeout_ms=20000, n_serial_waves=serial,
                           client_wg_waves=args.client_wg_waves)

    out = {"workload": "4096 accounts, 2-account transfers, 32-account audits, 90/10, "
                       f"Zipf theta; {args.sched_txns} txns per evaluation stream; "
                       "64 client waves (routed: 48 normal + 16 serialized-lane)"}
    with eng(0) as plain, eng(16) as routed:
        t0 = time.time()
        rep = sched.train(plain, n_txns=1 << 17, steps=300)
        cal = sched.calibrate(plain, routed)
        calk = sched.calibrate_heuristic(plain, routed)
        out["training"] = {"samples": rep.samples, "positive_rate": rep.positive_rate,
                           "loss_first": float(rep.loss[0]), "loss_last": float(rep.loss[-1]),
                           "loss_constant_predictor": rep.constant_loss, "steps": 300,
                           "tau": cal["tau"], "tau_per_level": cal["per_level"].levels,
                           "k_per_level": calk["per_level"].levels,
                           "seconds": time.time() - t0}
        log(f"scheduler: trained ({rep.samples} samples), tau={cal['tau']}")
        rows = {}
        for th in (0.6, 0.9, 1.2):
            bank = sched.c1_bank(args.sched_txns, seed=77, theta=th)
            sched.history(plain, sched.c1_bank(1 << 17, seed=76, theta=th))
            sched.transfer_state(plain, routed)
            sp = sched.run_logged(plain, bank)
            routed.bank_generate(bank)
            nh = sched.route(routed, cal["per_level"])
            routed.fill(INIT)
            sl = routed.bank_run()
            op, _ = routed.bank_read_stream()
            nk = sched.route_heuristic(routed, calk["per_level"])
            routed.fill(INIT)
            sk = routed.bank_run()
            routed.bank_route(((op & 0xFF) == 0).astype(np.uint8))
            routed.fill(INIT)
            sa = routed.bank_run()
            rows[str(th)] = {
                "plain": {"txn_per_s": sp["txn_per_s"], "abort_rate": sp["abort_rate"]},
                "learned": {"txn_per_s": sl["txn_per_s"], "abort_rate": sl["abort_rate"],
                            "routed": nh},
                "counter_rule": {"txn_per_s": sk["txn_per_s"], "abort_rate": sk["abort_rate"],
                                 "routed": nk},
                "all_transfers_routed": {"txn_per_s": sa["txn_per_s"],
                                         "abort_rate": sa["abort_rate"]},
                "learned_over_plain": sl["txn_per_s"] / sp["txn_per_s"],
            }
            log(f"scheduler theta={th}: plain {sp['txn_per_s']:.3g} learned "
                f"{sl['txn_per_s']:.3g} all {sa['txn_per_s']:.3g}")
        out["theta"] = rows
        params = plain.pred_get_params()
    return out, params, cal["per_level"]


def bounded_lane_section(csmv, args):
    """Where the routing decision matters: a serialized lane of a bounded size (2 of 64
    client waves), 2-account transfers and 32-account audits over 1M Zipf accounts.
    Routing every transfer is capped by the lane's two waves; arrival order aborts on the
    hot accounts; a router that sends only the contended transfers gets both.  Compared on
    fresh streams at each skew: arrival order, every transfer routed, the counter rule
    (route when an item's abort bucket exceeds k) and the learned predictor, both with
    thresholds calibrated per contention level on the training skews; the final balances
    of all four runs must be identical."""
    import numpy as np
    from csmv import sched

    S = args.bounded_serial_waves

    def mk(n, seed, th):
        return csmv.Bank(n_txns=n, seed=seed, n_reads=2, n_reads_ro=32, pct_ro=10, zipf=True,
                         theta=th)

    def eng(serial):
        return csmv.Engine(N_ACCOUNTS, depth=8, client_waves=64, wg_waves=6, variant="stamp",
                           timeout_ms=30000, client_wg_waves=1, n_serial_waves=serial)

    out = {"workload": f"1,048,576 Zipf accounts, 2-account transfers, 32-account audits, "
                       f"90/10, {args.bounded_txns} txns per evaluation stream; 64 client "
                       f"waves, {S} of them the serialized lane",
           "serial_lane_waves": S}
    with eng(0) as plain, eng(S) as routed:
        t0 = time.time()
        rep = sched.train(plain, n_txns=1 << 16, steps=300, make_bank=mk)
        cal = sched.calibrate(plain, routed, n_txns=1 << 16, make_bank=mk)
        calk = sched.calibrate_heuristic(plain, routed, n_txns=1 << 16, make_bank=mk)
        out["training"] = {"samples": rep.samples, "positive_rate": rep.positive_rate,
                           "loss_first": float(rep.loss[0]), "loss_last": float(rep.loss[-1]),
                           "tau_per_level": cal["per_level"].levels,
                           "k_per_level": calk["per_level"].levels,
                           "seconds": time.time() - t0}
        rows = {}
        for th in (0.6, 0.9, 1.2):
            bank = mk(args.bounded_txns, 77, th)
            sched.history(plain, mk(1 << 16, 76, th))
            sched.transfer_state(plain, routed)
            sp = sched.run_logged(plain, bank)
            finals = [plain.read_state()]
            routed.bank_generate(bank)
            op, _ = routed.bank_read_stream()
            res = {"plain": {"txn_per_s": sp["txn_per_s"], "abort_rate": sp["abort_rate"]}}
            for name, route in (
                    ("all_transfers_routed",
                     lambda: routed.bank_route(((op & 0xFF) == 0).astype(np.uint8)) or
                     int(((op & 0xFF) == 0).sum())),
                    ("counter_rule", lambda: sched.route_heuristic(routed, calk["per_level"])),
                    ("learned", lambda: sched.route(routed, cal["per_level"]))):
                nh = route()
                routed.fill(INIT)
                st = routed.bank_run()
                finals.append(routed.read_state())
                res[name] = {"txn_per_s": st["txn_per_s"], "abort_rate": st["abort_rate"],
                             "routed": nh}
            res["final_state_equal"] = all(np.array_equal(finals[0], f) for f in finals[1:])
            res["learned_over_best_other"] = res["learned"]["txn_per_s"] / max(
                res[k]["txn_per_s"] for k in ("plain", "all_transfers_routed", "counter_rule"))
            rows[str(th)] = res
            log(f"bounded lane theta={th}: plain {sp['txn_per_s']:.3g} all "
                f"{res['all_transfers_routed']['txn_per_s']:.3g} rule "
                f"{res['counter_rule']['txn_per_s']:.3g} learned {res['learned']['txn_per_s']:.3g}")
        out["theta"] = rows
    return out


def class_rate(arr, lat):
    """Requests and window of one class: first arrival to last commit (arrival +
    latency), in ticks."""
    import numpy as np
    end = (arr.astype(np.int64) + lat.astype(np.int64)).max()
    return arr.size, int(end - int(arr.min()))


def class_rate99(arr, lat, khz):
    """Requests per second of one class until 99 % of its requests have committed: 0.99 n
    over first arrival to the 99th-percentile commit time (a straggler tail, such as SETs
    queued behind the hottest key, does not set it)."""
    import numpy as np
    end = arr.astype(np.int64) + lat.astype(np.int64)
    t99 = float(np.percentile(end, 99)) - float(arr.min())
    return 0.99 * arr.size / (t99 / (khz * 1e3))


def serving_section(csmv, args, params, tau):
    """C5: offered-load sweep, with and without the predictor; the cache class split over
    args.serve_cache_parts commit servers by set."""
    import dataclasses

    import numpy as np
    from csmv import sched

    loads = [float(x) for x in args.serve_loads.split(",")]
    P = args.serve_cache_parts
    m = csmv.Memcd(n_reqs=args.serve_cache_reqs, seed=1)
    b = csmv.Bank(n_txns=args.serve_bank_txns, seed=7, n_reads=64, n_reads_min=2, pct_ro=0,
                  zipf=True, theta=0.99)

    def hist(seed):
        return csmv.Bank(n_txns=1 << 13, seed=seed, n_reads=64, n_reads_min=2, pct_ro=0,
                         zipf=True, theta=0.99)

    with csmv.Engine(m.n_items, depth=4, client_waves=1, wg_waves=6) as gen:
        gen.memcd_generate(m)
        key, op = gen.memcd_read_stream()
    parts = csmv.memcd_partition(key, op, m.n_sets, P)
    local = dataclasses.replace(m, n_sets=m.n_sets // P)
    res = {"workload": f"{m.n_reqs} cache requests (the first requests of the cache config "
                       f"stream, cold cache; 80 %) split by set over {P} commit servers of "
                       f"{args.serve_cache_waves} client waves each, + {b.n_txns} bank "
                       "transactions (transfer over the first 2 of 2..64 Zipf 0.99 reads of 1M "
                       "accounts; 20 %), Poisson arrivals; modes without_predictor / "
                       "with_predictor: the bank class on one commit server with 16 client "
                       "waves (8 serialized-lane when routed, escalation after "
                       f"{args.escalate_after} failed attempts); mode pipelined_bank_lane: "
                       "the bank class on the pipelined serialized lane, 4 commit server "
                       f"workgroups, {args.serve_pipe_waves} client waves",
           "cache_partitions": P,
           "pool_budget_bytes": args.serve_budget,
           "bank_escalate_after": args.escalate_after,
           "loads": {}}
    caches = [csmv.Engine(local.n_items, depth=4, client_waves=args.serve_cache_waves, wg_waves=6,
                          variant="stamp", timeout_ms=30000, client_wg_waves=args.client_wg_waves)
              for _ in range(P)]
    try:
        with csmv.Engine(N_ACCOUNTS, depth=8, client_waves=16, wg_waves=6, variant="stamp",
                         timeout_ms=30000, n_serial_waves=8,
                         client_wg_waves=args.serve_bank_client_wg_waves,
                         escalate_after=args.escalate_after) as eb:
            for ec, (k, o) in zip(caches, parts):
                ec.memcd_load(local, k, o)
            # the predictor (trained on the bank-plain abort logs) is fine-tuned on this
            # partition's own logs: history stream 1 gives the counters, stream 2 the labels
            t0 = time.time()
            eb.pred_set_params(params)
            for i, seed in enumerate((101, 102)):
                eb.fill(INIT)
                eb.bank_generate(hist(seed))
                csmv.serve(None, eb, offered_rps=loads[0], pct_cache=80, seed=3 + i)
                if i == 0:
                    eb.abort_fold(32)
                else:
                    eb.pred_clear_dataset()
                    n = eb.pred_collect()
                    loss = eb.pred_train(200, 1e-2)
                    eb.abort_fold(1)
            res["predictor"] = {"tau": getattr(tau, "levels", tau), "finetune_samples": n,
                                "loss_first": float(loss[0]), "loss_last": float(loss[-1]),
                                "seconds": time.time() - t0}
            eb.bank_generate(b)
            n_hot = sched.route(eb, tau)
            bop, _ = eb.bank_read_stream()
            bank_bytes = int((12 * ((bop >> 16) & 0xFF).astype(np.int64) + 36).sum())
            def measure(e_bank, mode, load, budget, routed):
                """One serving run at one offered load: the row of the JSON line."""
                for ec in caches:
                    ec.fill(0)
                e_bank.fill(INIT)
                out = csmv.serve(caches, e_bank, offered_rps=load, pct_cache=80, seed=1,
                                 pool_budget=budget)
                khz = out["wallclock_khz"]
                lc = np.concatenate([ec.serve_latency() for ec in caches])
                ac = np.concatenate([ec.serve_arrivals() for ec in caches])
                lb, ab = e_bank.serve_latency(), e_bank.serve_arrivals()
                nc_, wc = class_rate(ac, lc)
                nb_, wb = class_rate(ab, lb)
                _, wall = class_rate(np.concatenate([ac, ab]), np.concatenate([lc, lb]))
                # algorithmic bytes: cache requests as cache_bytes, bank transactions
                # 12 R + 12 W + 12 (R = 2..64 reads, W = 2)
                nbytes = bank_bytes + sum(cache_bytes(ec.memcd_read_stream()[1],
                                                      ec.memcd_read_log()["res"])
                                          for ec in caches)
                row = {"offered_rps": load, "kernel_ms": out["kernel_ms"],
                       "offered_window_ms": float(max(ac.max(), ab.max())) / khz,
                       "req_per_s": (nc_ + nb_) / (wall / (khz * 1e3)),
                       "cache_req_per_s": nc_ / (wc / (khz * 1e3)),
                       "bank_txn_per_s": nb_ / (wb / (khz * 1e3)),
                       "req_per_s_99": class_rate99(np.concatenate([ac, ab]),
                                                    np.concatenate([lc, lb]), khz),
                       "cache_req_per_s_99": class_rate99(ac, lc, khz),
                       "bank_txn_per_s_99": class_rate99(ab, lb, khz),
                       "cache_offered_rps": 0.8 * load, "bank_offered_rps": 0.2 * load,
                       "all": pct_us(np.concatenate([lc, lb]), khz), "cache": pct_us(lc, khz),
                       "bank": pct_us(lb, khz),
                       "cache_abort_rate": out["cache"]["abort_rate"],
                       "bank_abort_rate": out["bank"]["abort_rate"],
                       "bank_routed": routed,
                       "bank_escalated": out["bank"]["escalated"],
                       "admit_stalls": out["admit_stalls"],
                       "pool_budget_bytes": budget,
                       "peak_inflight_bytes": out["peak_inflight"],
                       "roofline": roofline(nbytes, out["kernel_ms"])}
                res["loads"].setdefault(mode, []).append(row)
                log(f"serving {mode} {load:.3g} req/s: {row['req_per_s']:.3g} served (cache "
                    f"{row['cache_req_per_s']:.3g} / {row['cache_req_per_s_99']:.3g} to 99 %, "
                    f"bank {row['bank_txn_per_s']:.3g}), p99 "
                    f"{row['all']['p99_us']:.1f} us (cache {row['cache']['p99_us']:.1f}, bank "
                    f"{row['bank']['p99_us']:.1f}), stalls {row['admit_stalls']}")

            for mode in ("without_predictor", "with_predictor"):
                for load in loads:
                    if mode == "without_predictor":
                        eb.bank_route(None)
                    else:
                        sched.route(eb, tau)
                    measure(eb, mode, load, args.serve_budget,
                            n_hot if mode == "with_predictor" else 0)
        # the bank class on the pipelined serialized lane: every bank transaction executed
        # by the commit servers in commit order (nothing aborts, no routing), its accounts
        # split over 4 commit server workgroups; admission engages at overload only
        with csmv.Engine(N_ACCOUNTS, depth=8, client_waves=args.serve_pipe_waves, wg_waves=6,
                         variant="stamp_pipe", timeout_ms=30000, commit_parts=4,
                         client_wg_waves=1) as ep:
            ep.bank_generate(b)
            for load in loads:
                measure(ep, "pipelined_bank_lane", load, args.serve_pipe_budget, 0)
    finally:
        for ec in caches:
            ec.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--variant", default="stamp_pipe",
                    help="the headline's commit path: stamp_pipe (the pipelined serialized "
                         "lane) or stamp (one committer wave)")
    ap.add_argument("--waves", type=int, default=int(os.environ.get("CSMV_BENCH_WAVES", "248")))
    ap.add_argument("--serial-waves", type=int, default=247)
    ap.add_argument("--wg-waves", type=int, default=6,
                    help="waves per workgroup of the one-committer server: the committer, the "
                         "clock advancer and wg_waves - 2 receivers")
    ap.add_argument("--pipe-wg-waves", type=int, default=16,
                    help="waves per workgroup of each pipelined commit server: w / 4 "
                         "receivers (1 below 8 waves), one publisher, the rest committers")
    ap.add_argument("--commit-parts", type=int, default=8,
                    help="commit server workgroups of the pipelined lane (items split by hash)")
    ap.add_argument("--one-waves", type=int, default=64,
                    help="client waves of the one-committer line (and the reference variants)")
    ap.add_argument("--one-serial-waves", type=int, default=14,
                    help="serialized-lane waves of the one-committer line")
    ap.add_argument("--client-wg-waves", type=int, default=1,
                    help="client waves per client workgroup (one per CU spreads the clients' "
                         "memory traffic over more CUs)")
    ap.add_argument("--serve-bank-client-wg-waves", type=int, default=0,
                    help="client waves per workgroup of the serving loop's bank partition "
                         "(0: six)")
    ap.add_argument("--stage-ahead", type=int, default=5,
                    help="requests the receivers stage ahead of the committer")
    ap.add_argument("--router", default="all", choices=("none", "transfers", "all"),
                    help="none: every transaction on the normal (client) lane; transfers: "
                         "every transfer on the serialized lane; all: audits too")
    ap.add_argument("--txns", type=int, default=1 << 24)
    ap.add_argument("--variant-txns", type=int, default=1 << 16,
                    help="stream length for the side-by-side reference variants")
    ap.add_argument("--memcd-reqs", type=int, default=1 << 24)
    ap.add_argument("--memcd-waves", type=int, default=6)
    ap.add_argument("--sched-txns", type=int, default=1 << 18)
    ap.add_argument("--serve-loads", default="5e6,50e6,100e6,200e6,500e6")
    ap.add_argument("--serve-cache-reqs", type=int, default=1 << 18)
    ap.add_argument("--serve-bank-txns", type=int, default=1 << 16)
    ap.add_argument("--serve-cache-waves", type=int, default=12,
                    help="client waves of each cache partition of the serving loop")
    ap.add_argument("--serve-cache-parts", type=int, default=4,
                    help="commit servers the serving loop's cache class is split over (by set)")
    ap.add_argument("--serve-budget", type=int, default=65536,
                    help="admission cap on in-flight version-pool bytes of the serving loop")
    ap.add_argument("--serve-pipe-waves", type=int, default=48,
                    help="client waves of the serving loop's bank class on the pipelined lane")
    ap.add_argument("--serve-pipe-budget", type=int, default=196608,
                    help="admission cap of the serving runs with the pipelined bank lane")
    ap.add_argument("--serial-waves-all", type=int, default=247,
                    help="serialized-lane waves of the run with every transaction routed")
    ap.add_argument("--serial-waves-transfers", type=int, default=216,
                    help="serialized-lane waves of the run with only the transfers routed (the "
                         "other client waves run the audits on the normal lane)")
    ap.add_argument("--no-one", action="store_true",
                    help="skip the one-committer serialized-lane line")
    ap.add_argument("--escalate-after", type=int, default=2,
                    help="failed attempts after which the stamp_escalate variant line sends a "
                         "transaction on the serialized lane")
    ap.add_argument("--no-route-all", action="store_true",
                    help="skip the line with the other router (transfers only when the "
                         "headline routes every transaction, every transaction otherwise)")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--no-memcd", action="store_true")
    ap.add_argument("--no-sched", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=1 << 22)
    ap.add_argument("--cpu-sample-1t", type=int, default=1 << 20,
                    help="transactions of the single-client CPU run")
    ap.add_argument("--plain-txns", type=int, default=1 << 20)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-bounded", action="store_true")
    ap.add_argument("--bounded-txns", type=int, default=1 << 18)
    ap.add_argument("--bounded-serial-waves", type=int, default=2)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              f"(the CSMV commit path has one global clock per engine)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import csmv

    bank = csmv.Bank(n_txns=args.txns, seed=1, n_reads=4, n_reads_ro=4, pct_ro=10, zipf=True,
                     theta=0.99)

    def make(variant, b, router, serial, escalate=0, waves=None):
        pipe = variant == "stamp_pipe"
        e = csmv.Engine(N_ACCOUNTS, depth=8, atr_entries=4096,
                        client_waves=waves or (args.waves if pipe else args.one_waves),
                        wg_waves=args.pipe_wg_waves if pipe else args.wg_waves, variant=variant,
                        timeout_ms=8000, n_serial_waves=serial if router != "none" else 0,
                        stage_ahead=args.stage_ahead, client_wg_waves=args.client_wg_waves,
                        escalate_after=escalate, commit_parts=args.commit_parts if pipe else 0)
        e.bank_generate(b)
        if router == "transfers":
            op, _ = e.bank_read_stream()
            e.bank_route(((op & 0xFF) == 0).astype(np.uint8))
        elif router == "all":
            e.bank_route(np.ones(b.n_txns, dtype=np.uint8))
        return e

    pipe_h = args.variant == "stamp_pipe"
    waves_h = args.waves if pipe_h else args.one_waves
    wg_h = args.pipe_wg_waves if pipe_h else args.wg_waves
    # pipelined server workgroup of w waves: w / 4 receivers (1 below 8 waves), a publisher,
    # the rest committers
    committers = (wg_h - 1 - (wg_h // 4 if wg_h >= 8 else 1)) * args.commit_parts if pipe_h else 1
    log(f"headline: {args.txns} txns, router {args.router}")
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
    log(f"headline: {value:.4g} txn/s")

    alg_bytes = bytes_per_commit(runs[-1])
    achieved = alg_bytes / (runs[-1]["kernel_ms"] * 1e-3) / 1e9
    traffic = load_traffic(runs[-1]["commits"], "csmv::bank_pipe_kernel<4, 2>" if pipe_h
                           else "csmv::bank_kernel<4, 2>")
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic}
    # the per-lane protocol spends at least two same-address atomics per committed
    # update (ready bit + commit counter); its ceiling is the measured rate / 2
    plain_ceiling = atomic_rate / 2.0
    route_all, route_transfers = {}, {}
    if not args.no_route_all and args.router != "all":
        # the same stream with audits on the serialized lane too: no transaction aborts
        ea = make(args.variant, bank, "all", args.serial_waves_all)
        ea.fill(INIT)
        ra = ea.bank_run()
        ea.close()
        route_all = {"txn_per_s": ra["txn_per_s"], "abort_rate": ra["abort_rate"],
                     "kernel_ms": ra["kernel_ms"], "serial_lane_waves": args.serial_waves_all,
                     "commits_ro": ra["commits_ro"]}
        log(f"every transaction on the serialized lane: {ra['txn_per_s']:.4g} txn/s")
    elif not args.no_route_all:
        # the same stream with only the transfers on the serialized lane: audits run on the
        # normal lane as multi-version snapshot reads and abort when a version they need has
        # left its account's 8-deep ring
        et = make(args.variant, bank, "transfers", args.serial_waves_transfers)
        et.fill(INIT)
        rt = et.bank_run()
        et.close()
        route_transfers = {"txn_per_s": rt["txn_per_s"], "abort_rate": rt["abort_rate"],
                           "kernel_ms": rt["kernel_ms"], "serial_lane_waves": args.serial_waves_transfers,
                           "aborts_dataage": rt["aborts_dataage"], "commits_ro": rt["commits_ro"]}
        log(f"transfers on the serialized lane, audits on the normal lane: {rt['txn_per_s']:.4g} txn/s")
    one = {}
    if not args.no_one and args.variant == "stamp_pipe":
        # the same stream on the serialized lane with one committer wave, whose clients install
        # their own versions (one client-side install hand-off per request on a hot account)
        eo = make("stamp", bank, args.router, args.one_serial_waves)
        eo.fill(INIT)
        eo.bank_run()
        eo.fill(INIT)
        ro1 = eo.bank_run()
        eo.close()
        one = {"txn_per_s": ro1["txn_per_s"], "abort_rate": ro1["abort_rate"],
               "kernel_ms": ro1["kernel_ms"], "client_waves": args.one_waves,
               "serial_lane_waves": args.one_serial_waves, "server_wg_waves": args.wg_waves,
               "committer_busy_frac": ro1["prof"]["commit_busy"] / (ro1["kernel_ms"] * 1e6)}
        log(f"one committer wave: {ro1['txn_per_s']:.4g} txn/s")
    variants = {}
    if not args.no_variants:
        vbank = csmv.Bank(n_txns=args.variant_txns, seed=1, n_reads=4, n_reads_ro=4, pct_ro=10,
                          zipf=True, theta=0.99)
        for v in REF_VARIANTS + ("stamp", "stamp_escalate"):
            log(f"variant {v}")
            esc = args.escalate_after if v == "stamp_escalate" else 0
            ev = make("stamp" if esc else v, vbank, "none", 0, escalate=esc)
            ev.fill(INIT)
            r = ev.bank_run()
            ev.close()
            variants[v] = {"txn_per_s": r["txn_per_s"], "abort_rate": r["abort_rate"],
                           "kernel_ms": r["kernel_ms"], "txns": args.variant_txns,
                           "lane": "normal (client-executed CSMV)"}
            if esc:
                variants[v]["escalate_after"] = esc
                variants[v]["escalated"] = r["escalated"]
                variants[v]["lane"] = ("normal, transactions failing " + str(esc) +
                                       " attempts executed on the serialized lane")
    # the target's measure: the headline over the plain per-lane protocol's ceiling (the
    # measured same-address atomic rate / 2 atomics per committed update)
    atomics = {"same_address_per_s": atomic_rate,
               "plain_per_lane_ceiling_txn_per_s": plain_ceiling,
               "value_over_plain_ceiling": value / plain_ceiling}
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
            "commit_path": {
                ("stamp_pipe", "all"): "pipelined serialized lane for every transaction: client "
                    "waves hold four transactions a lane and stage each request's (up to 256 "
                    "transactions) producer chains; committer waves apply each account's "
                    "request-wide delta with one ordered 64-bit atomic, whose return gives every "
                    "read its value in the serial order, and install the versions",
                ("stamp_pipe", "transfers"): "pipelined serialized lane for every transfer; "
                    "audits on the normal lane as multi-version snapshot reads",
            }.get((args.variant, args.router), args.variant),
            "router": args.router,
            "client_waves": waves_h,
            "serial_lane_waves": args.serial_waves if args.router != "none" else 0,
            "server_wg_waves": wg_h,
            "commit_parts": args.commit_parts if pipe_h else 1,
            "committer_waves": committers,
            "client_wg_waves": args.client_wg_waves,
            "stage_ahead": args.stage_ahead,
            "client_lanes": waves_h * 64,
            "txns_per_lane": 4 if pipe_h else 1,
            "record_ring": 4096,
        },
        "aborts_by_class": {k: sum(r[k] for r in runs) for k in
                            ("aborts_dataage", "aborts_readwrite", "aborts_prevalid",
                             "aborts_recordage", "aborts_unknown")},
        # busy time of the committer waves over the kernel time, per committer wave
        "committer_busy_frac": sum(r["prof"]["commit_busy"] for r in runs) / (kernel_s * 1e9) / committers,
        "abort_rate": aborts / (aborts + commits),
        "wall_s_timed_steps": wall,
        "roofline": roofline,
        "atomics": atomics,
        "every_transaction_serialized": route_all,
        "audits_on_normal_lane": route_transfers,
        "one_committer_serialized_lane": one,
        "variants": variants,
    }
    if not args.no_plain:
        out["bank_plain"] = bank_plain_section(csmv, args)
    if not args.no_memcd:
        out["memcd"] = memcd_section(csmv, args)
    params, tau = None, 0.002
    if not args.no_sched:
        out["scheduler"], params, tau = scheduler_section(csmv, args)
    if not args.no_bounded:
        out["scheduler_bounded_lane"] = bounded_lane_section(csmv, args)
    if not args.no_serve:
        if params is None:
            with csmv.Engine(4096, depth=4, client_waves=64, wg_waves=8, variant="stamp",
                             client_wg_waves=args.client_wg_waves) as pe:
                from csmv import sched
                sched.train(pe, n_txns=1 << 17, steps=300)
                params = pe.pred_get_params()
        out["serving"] = serving_section(csmv, args, params, tau)
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(op, addr, args.txns, min(args.cpu_sample, args.txns),
                                           min(args.cpu_sample_1t, args.txns))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
