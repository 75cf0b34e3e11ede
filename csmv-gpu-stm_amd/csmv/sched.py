"""Learned contention scheduler over the engine's predictor (csrc/mlp.hpp).

Training recipe: for each Zipf skew, run a history stream in arrival order and fold
its abort log into the engine's per-bucket counters, then run a second stream and
collect its transfers' features (under the history counters) with their labels
(aborted at least once).  The predictor is trained on the union by the engine's own
forward / backward / Adam kernels.

At run time the scheduler folds the latest abort log, predicts every transfer of the
resident stream, and partitions the stream: predicted-hot transfers are grouped into
serialized-lane wavefronts (executed by the commit server in commit order), the rest
keep their arrival order on the normal CSMV lane.  The predictor never changes commit
semantics: every transaction still commits through CSMV, and arrival-order CSMV
(no routing) is the fallback.

No counterpart in the reference, whose drivers dispatch in arrival order
(bank/csmv_sh/exe.cu:142-269).
"""
from __future__ import annotations

import dataclasses
from typing import Iterable, Optional

import numpy as np

from . import Bank, Engine

INIT = 1000


@dataclasses.dataclass
class TrainReport:
    samples: int
    positive_rate: float
    loss: np.ndarray
    thetas: tuple

    @property
    def constant_loss(self) -> float:
        """Cross-entropy of the best constant predictor (the label entropy)."""
        q = min(max(self.positive_rate, 1e-12), 1 - 1e-12)
        return float(-(q * np.log(q) + (1 - q) * np.log(1 - q)))


def c1_bank(n_txns: int, seed: int, theta: Optional[float]) -> Bank:
    """The 'bank plain' shape (2-account transfers, 32-account audits, 90/10), with
    account choice Zipf(theta) when theta is given."""
    return Bank(n_txns=n_txns, seed=seed, n_reads=2, n_reads_ro=32, pct_ro=10,
                zipf=theta is not None, theta=theta if theta is not None else 0.99)


def run_logged(e: Engine, bank: Bank, init: int = INIT) -> dict:
    e.bank_generate(bank)
    e.fill(init)
    return e.bank_run()


def history(e: Engine, bank: Bank, init: int = INIT, decay_shift: int = 32) -> dict:
    """Runs bank in arrival order and folds its abort log into the counters."""
    st = run_logged(e, bank, init)
    e.abort_fold(decay_shift)
    return st


def train(e: Engine, thetas: Iterable[float] = (0.6, 0.9, 1.2), n_txns: int = 1 << 17,
          steps: int = 300, lr: float = 1e-2, seed: int = 1000, init: int = INIT,
          make_bank=c1_bank) -> TrainReport:
    """Trains e's predictor on its own abort logs (e runs in arrival order)."""
    thetas = tuple(thetas)
    e.pred_init(seed)
    e.pred_clear_dataset()
    n = 0
    for i, th in enumerate(thetas):
        history(e, make_bank(n_txns, seed + 2 * i, th), init)
        run_logged(e, make_bank(n_txns, seed + 2 * i + 1, th), init)
        n = e.pred_collect()
    _, y = e.pred_dataset(n)
    loss = e.pred_train(steps, lr)
    return TrainReport(samples=n, positive_rate=float(y.mean()), loss=loss, thetas=thetas)


def transfer_state(src: Engine, dst: Engine):
    """Copies the trained parameters and the abort counters to another engine."""
    dst.pred_set_params(src.pred_get_params())
    dst.abort_load_counters(src.abort_counters())


def contention_level(counters: np.ndarray) -> float:
    """How contended the folded history was: the mean floor(log2(1 + A)) over the abort
    buckets (0 for a history without aborts)."""
    a = np.asarray(counters, dtype=np.float64)
    return float(np.mean(np.floor(np.log2(1.0 + a))))


@dataclasses.dataclass
class Thresholds:
    """A routing threshold per contention level seen in calibration: at run time the level
    of the engine's current abort counters picks the threshold of the nearest calibrated
    level, so one scheduler serves mild and heavy contention."""
    levels: list   # [(contention level, threshold)], ascending by level

    def pick(self, counters: np.ndarray):
        lv = contention_level(counters)
        return min(self.levels, key=lambda t: abs(t[0] - lv))[1]


def _threshold(e: Engine, thr):
    return thr.pick(e.abort_counters()) if isinstance(thr, Thresholds) else thr


def route(e: Engine, tau=0.5) -> int:
    """Learned routing of e's resident stream (tau: a threshold, or Thresholds chosen by
    the current abort counters); returns the serialized-lane count."""
    return e.bank_route_learned(_threshold(e, tau))


# ---- the non-learned rule the predictor is compared with -------------------------------------
# Route a transfer when the abort counter of any of its items' buckets exceeds k (the same
# counters and bucket hash the predictor's features use, csrc/mlp.hpp).
KS = (0, 1, 2, 4, 8, 16, 32, 64, 128, 256, 1024)
FEAT_ITEMS = 4


def _bucket_hash(x: np.ndarray) -> np.ndarray:
    """The device's 32-bit finaliser (feat_mix in csrc/mlp.hpp), vectorised."""
    x = x.astype(np.uint64)
    m = np.uint64(0xFFFFFFFF)
    x ^= x >> np.uint64(16)
    x = (x * np.uint64(0x7FEB352D)) & m
    x ^= x >> np.uint64(15)
    x = (x * np.uint64(0x846CA68B)) & m
    x ^= x >> np.uint64(16)
    return x


def heuristic_flags(op: np.ndarray, addr: np.ndarray, counters: np.ndarray, k: int) -> np.ndarray:
    """1 for each transfer with an item whose abort bucket holds more than k."""
    T = op.size
    A = addr.reshape(-1, T)
    nr = (op >> 16) & 0xFF
    worst = np.zeros(T, dtype=np.uint64)
    for j in range(min(FEAT_ITEMS, A.shape[0])):
        c = counters[(_bucket_hash(A[j]) & np.uint64(counters.size - 1)).astype(np.int64)]
        worst = np.maximum(worst, np.where(j < nr, c.astype(np.uint64), np.uint64(0)))
    return (((op & 0xFF) == 0) & (worst > np.uint64(k))).astype(np.uint8)


def route_heuristic(e: Engine, k) -> int:
    """Routes e's resident stream by the counter rule (k: a bound, or Thresholds chosen by
    the current abort counters); returns the serialized-lane count."""
    op, addr = e.bank_read_stream()
    f = heuristic_flags(op, addr, e.abort_counters(), _threshold(e, k))
    e.bank_route(f)
    return int(f.sum())


# ---- threshold calibration on the training skews -------------------------------------------
TAUS = (0.002, 0.005, 0.01, 0.02, 0.05, 0.1, 0.2, 0.3, 0.5, 0.7)


def _calibrate(plain: Engine, routed: Engine, apply, candidates, thetas, n_txns, seed, init,
               make_bank):
    """For each skew a history stream refreshes the counters (plain engine, arrival order),
    then a validation stream is routed with every candidate threshold on the routed engine.
    Returns the candidate with the highest geometric-mean committed txns/s over the skews,
    the best candidate per skew keyed by the history's contention level, and the table."""
    thetas, candidates = tuple(thetas), tuple(candidates)
    logs = {c: [] for c in candidates}
    table, levels = [], []
    for i, th in enumerate(thetas):
        history(plain, make_bank(n_txns, seed + 2 * i, th), init)
        transfer_state(plain, routed)
        lvl = contention_level(routed.abort_counters())
        routed.bank_generate(make_bank(n_txns, seed + 2 * i + 1, th))
        best = None
        for c in candidates:
            nh = apply(routed, c)
            routed.fill(init)
            st = routed.bank_run()
            logs[c].append(np.log(max(st["txn_per_s"], 1.0)))
            table.append({"theta": th, "level": lvl, "threshold": c, "routed": nh,
                          "txn_per_s": st["txn_per_s"], "abort_rate": st["abort_rate"]})
            if best is None or st["txn_per_s"] > best[1]:
                best = (c, st["txn_per_s"])
        levels.append((lvl, best[0]))
    overall = max(candidates, key=lambda c: float(np.mean(logs[c])))
    return overall, Thresholds(sorted(levels)), table


def calibrate(plain: Engine, routed: Engine, thetas: Iterable[float] = (0.6, 0.9, 1.2),
              taus: Iterable[float] = TAUS, n_txns: int = 1 << 17, seed: int = 2000,
              init: int = INIT, make_bank=c1_bank) -> dict:
    """Chooses the predictor's routing threshold on the training skews (fresh streams):
    "tau" is the one threshold with the highest geometric-mean committed txns/s,
    "per_level" the best threshold of each skew keyed by its contention level."""
    overall, per_level, table = _calibrate(plain, routed, lambda e, t: e.bank_route_learned(t),
                                           taus, thetas, n_txns, seed, init, make_bank)
    return {"tau": overall, "per_level": per_level, "table": table}


def calibrate_heuristic(plain: Engine, routed: Engine, thetas: Iterable[float] = (0.6, 0.9, 1.2),
                        ks: Iterable[int] = KS, n_txns: int = 1 << 17, seed: int = 2000,
                        init: int = INIT, make_bank=c1_bank) -> dict:
    """The counter rule's bound chosen exactly like the predictor's threshold (calibrate)."""
    overall, per_level, table = _calibrate(plain, routed, route_heuristic, ks, thetas, n_txns,
                                           seed, init, make_bank)
    return {"k": overall, "per_level": per_level, "table": table}
