"""GPU tests of the learned contention predictor: abort-log folding, features, forward
and backward kernels against the CPU oracle, Adam, routing, and the scheduler's gain
over arrival-order CSMV at equal, bit-exact final state.

Tolerances (floating point): probabilities within 1e-4 absolute; gradients within
2e-3 of the largest gradient magnitude (float32 accumulation order and bf16 rounding of
dA at ties, against the float64 oracle that rounds dA the same way); loss within 1e-5
relative.  Counters, features, labels and routing lists are integer / exact: bit-exact.
"""
import numpy as np
import pytest

import csmv
import mlp_oracle as M
import orc
from csmv import sched

pytestmark = pytest.mark.gpu

INIT = 1000
N = 4096


def engine(waves=64, serial=0):
    return csmv.Engine(N, depth=4, atr_entries=4096, client_waves=waves, wg_waves=8,
                       variant="stamp", timeout_ms=8000, n_serial_waves=serial)


def check_replay(e, bank, stats, n=N):
    op, addr = e.bank_read_stream()
    ts, v, ab = e.bank_read_log()
    final = e.read_state()
    spec = orc.bank_spec(n, bank.n_txns, seed=bank.seed, n_reads=bank.n_reads,
                         n_reads_ro=bank.n_reads_ro, pct_ro=bank.pct_ro, zipf=bank.zipf,
                         theta=bank.theta, amount_max=bank.amount_max)
    rep, replayed = orc.bank_replay(spec, op, addr, INIT, ts, v, final)
    assert rep["fails"] == 0, rep
    assert rep["total"] == n * INIT
    assert stats["commits"] == bank.n_txns
    serial_state, _, _ = orc.bank_serial(spec, op, addr, INIT)
    assert np.array_equal(final, serial_state), "final state differs from arrival-order replay"
    return final


def bf16_exact_params(rng, scale=0.3):
    p = rng.uniform(-scale, scale, M.N_PARAMS).astype(np.float32)
    p[:M.B1] = M.bf16_value(M.bf16_bits(p[:M.B1]))
    return p


def test_fold_features_labels_match_oracle():
    bank = sched.c1_bank(1 << 16, seed=21, theta=0.9)
    print(f"\n[case] fold/features C1 shape theta=0.9 txns={bank.n_txns}", flush=True)
    with engine() as e:
        e.pred_init(1)
        st = sched.run_logged(e, bank)
        check_replay(e, bank, st)
        op, addr = e.bank_read_stream()
        _, _, ab = e.bank_read_log()
        rng = np.random.default_rng(9)
        prev = rng.integers(0, 5000, M.ABORT_BUCKETS).astype(np.uint32)
        e.abort_load_counters(prev)
        e.abort_fold(1)
        got = e.abort_counters()
        want = M.abort_fold(op, addr, bank.n_txns, ab, prev, decay_shift=1)
        assert np.array_equal(got, want)
        assert st["aborts"] > 0 and got.sum() > prev.sum() // 2
        n = e.pred_collect()
        X, y = e.pred_dataset(n)
    idx = np.nonzero((op & 0xFF) == 0)[0]
    assert n == idx.size
    Xo = M.features(op, addr, bank.n_txns, idx, want)
    assert np.array_equal(X, M.bf16_bits(Xo))
    assert np.array_equal(y, (ab[idx] > 0).astype(np.float32))


def test_forward_and_gradient_match_oracle():
    rng = np.random.default_rng(10)
    print("\n[case] forward/backward kernels vs oracle, n = 1000 and 4097 rows", flush=True)
    with engine(waves=8) as e:
        for n in (1000, 4097):
            X = (rng.integers(0, 40, (n, 64)) * (rng.random((n, 64)) < 0.15)) / 8
            Xb = M.bf16_bits(X)
            y = (rng.random(n) < 0.35).astype(np.float32)
            p = bf16_exact_params(rng)
            e.pred_set_params(p)
            assert np.array_equal(e.pred_get_params(), p)
            prob = e.pred_forward(Xb)
            _, po, _, _ = M.forward(X, p)
            assert np.max(np.abs(prob - po)) < 1e-4
            g, loss = e.pred_grad(Xb, y)
            go, lo = M.grad(X, y, p, round_w1=True, round_da=True)
            scale = np.max(np.abs(go))
            err = np.abs(g - go)
            print(f"[grad] n={n} max|g|={scale:.3e} max err={err.max():.3e} "
                  f"loss {loss:.6f} vs {lo:.6f}", flush=True)
            assert err.max() <= 2e-3 * scale
            for lo_i, hi_i in ((0, M.B1), (M.B1, M.W2), (M.W2, M.B2), (M.B2, M.N_PARAMS)):
                blk = np.max(np.abs(go[lo_i:hi_i]))
                assert np.max(err[lo_i:hi_i]) <= 2e-3 * max(blk, 1e-3 * scale)
            assert abs(loss - lo) <= 1e-5 * abs(lo)


def test_adam_first_step_and_training_reduce_loss():
    print("\n[case] Adam first step closed form; training on C1-shape abort logs", flush=True)
    with engine() as e:
        rep = sched.train(e, thetas=(0.6, 0.9, 1.2), n_txns=1 << 15, steps=1, lr=1e-2)
        # the first step from zero moments moves each parameter by lr against sign(g)
        e.pred_init(1000)
        p0 = e.pred_get_params()
        X, y = e.pred_dataset(rep.samples)
        g, _ = e.pred_grad(X, y)
        e.pred_train(1, 1e-2)
        p1 = e.pred_get_params()
        big = np.abs(g) > 1e-5
        assert big.sum() > 1000
        assert np.allclose((p1 - p0)[big], -1e-2 * np.sign(g[big]), rtol=0, atol=1e-5)
        assert np.all(np.abs(p1 - p0)[~big] <= 1e-2 + 1e-6)
        loss = e.pred_train(300, 1e-2)
        print(f"[train] samples={rep.samples} positives={rep.positive_rate:.3f} loss "
              f"{loss[0]:.4f} -> {loss[-1]:.4f} (constant predictor {rep.constant_loss:.4f})",
              flush=True)
        assert loss[-1] < loss[0]
        assert loss[-1] < rep.constant_loss
        # the device loss is the oracle's loss of the device parameters
        _, lo = M.grad(M.bf16_value(X).astype(np.float64), y, e.pred_get_params())
        _, lg = e.pred_grad(X, y)
        assert abs(lg - lo) <= 1e-5 * lo


def test_route_learned_matches_oracle_partition():
    bank = sched.c1_bank(1 << 16, seed=31, theta=1.2)
    print(f"\n[case] learned routing, theta=1.2 txns={bank.n_txns}", flush=True)
    with engine(serial=16) as e:
        sched.train(e, thetas=(1.2,), n_txns=1 << 15, steps=100)
        sched.history(e, sched.c1_bank(1 << 15, seed=30, theta=1.2))
        e.bank_generate(bank)
        p = e.pred_infer(0.5)
        nh = e.bank_route_learned(0.5)
        norm, hot = e.bank_read_route()
        op, _ = e.bank_read_stream()
        on, oh = M.route(op, p, 0.5)
        assert np.array_equal(norm, on) and np.array_equal(hot, oh) and nh == oh.size
        assert 0 < nh < bank.n_txns
        assert np.all(p[(op & 0xFF) != 0] == 0)
        e.fill(INIT)
        st = e.bank_run()
        check_replay(e, bank, st)


@pytest.fixture(scope="module")
def trained():
    """Predictor trained on abort logs of C1-shape runs at theta in {0.6, 0.9, 1.2}, with
    its routing threshold calibrated on the same skews (fresh streams)."""
    print("\n[case] training + threshold calibration, C1 shape", flush=True)
    plain, learned = engine(), engine(serial=16)
    rep = sched.train(plain, n_txns=1 << 17, steps=300)
    cal = sched.calibrate(plain, learned)
    print(f"[train] samples={rep.samples} loss {rep.loss[0]:.4f} -> {rep.loss[-1]:.4f} "
          f"(constant {rep.constant_loss:.4f}); tau={cal['tau']}", flush=True)
    yield plain, learned, rep, cal
    plain.close()
    learned.close()


@pytest.mark.parametrize("theta", [0.6, 0.9, 1.2])
def test_learned_scheduler_beats_arrival_order(trained, theta):
    """Evaluated on a fresh stream (seed not used in training): committed txns/s and
    abort rate of learned routing vs arrival-order CSMV, same bit-exact final state."""
    plain, learned, rep, cal = trained
    T = 1 << 18
    bank = sched.c1_bank(T, seed=77, theta=theta)
    print(f"\n[case] scheduler gain C1 shape theta={theta} txns={T}", flush=True)
    sched.history(plain, sched.c1_bank(1 << 17, seed=76, theta=theta))
    sched.transfer_state(plain, learned)
    st_plain = sched.run_logged(plain, bank)
    f_plain = check_replay(plain, bank, st_plain)
    learned.bank_generate(bank)
    nh = sched.route(learned, cal["tau"])
    learned.fill(INIT)
    st_l = learned.bank_run()
    f_l = check_replay(learned, bank, st_l)
    assert np.array_equal(f_plain, f_l)
    print(f"[gain] theta={theta} tau={cal['tau']} routed {nh}/{T}: plain "
          f"{st_plain['txn_per_s'] / 1e6:.2f}M txn/s abort {st_plain['abort_rate']:.3f} | "
          f"learned {st_l['txn_per_s'] / 1e6:.2f}M txn/s abort {st_l['abort_rate']:.3f}",
          flush=True)
    assert st_l["txn_per_s"] > 2 * st_plain["txn_per_s"]
    assert st_l["abort_rate"] < st_plain["abort_rate"]


def test_learned_router_beats_route_all_and_counter_rule_on_bounded_lane():
    """Where the routing decision matters: 2 of 64 client waves form the serialized lane,
    2-account transfers over 1M Zipf accounts.  Routing every transfer is capped by the
    lane; the counter rule (route when an item's abort bucket exceeds k) and the learned
    predictor get thresholds calibrated per contention level on the training skews.  On
    fresh streams the learned router commits more txns/s than arrival order, than routing
    every transfer and than the counter rule at one skew at least, with identical final
    balances in every run."""
    n_acc = 1 << 20

    def mk(n, seed, th):
        return csmv.Bank(n_txns=n, seed=seed, n_reads=2, n_reads_ro=32, pct_ro=10, zipf=True,
                         theta=th)

    def eng(serial):
        return csmv.Engine(n_acc, depth=8, client_waves=64, wg_waves=6, variant="stamp",
                           timeout_ms=30000, client_wg_waves=1, n_serial_waves=serial)

    print("\n[case] bounded serialized lane: training + per-level calibration", flush=True)
    wins = []
    with eng(0) as plain, eng(2) as routed:
        sched.train(plain, n_txns=1 << 16, steps=300, make_bank=mk)
        cal = sched.calibrate(plain, routed, n_txns=1 << 16, make_bank=mk)
        calk = sched.calibrate_heuristic(plain, routed, n_txns=1 << 16, make_bank=mk)
        for th in (0.6, 0.9, 1.2):
            bank = mk(1 << 17, 77, th)
            print(f"[case] bounded lane theta={th}", flush=True)
            sched.history(plain, mk(1 << 16, 76, th))
            sched.transfer_state(plain, routed)
            rate = {"plain": sched.run_logged(plain, bank)["txn_per_s"]}
            final = plain.read_state()
            routed.bank_generate(bank)
            op, _ = routed.bank_read_stream()
            for name in ("all", "rule", "learned"):
                if name == "all":
                    routed.bank_route(((op & 0xFF) == 0).astype(np.uint8))
                elif name == "rule":
                    sched.route_heuristic(routed, calk["per_level"])
                else:
                    sched.route(routed, cal["per_level"])
                routed.fill(INIT)
                st = routed.bank_run()
                check_replay(routed, bank, st, n_acc)
                assert np.array_equal(routed.read_state(), final)
                rate[name] = st["txn_per_s"]
            print(f"[rates] theta={th} " + " ".join(f"{k} {v / 1e6:.2f}M" for k, v in rate.items()),
                  flush=True)
            wins.append(rate["learned"] > max(rate["plain"], rate["all"], rate["rule"]))
    assert any(wins), wins
