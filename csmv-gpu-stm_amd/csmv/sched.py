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


def route(e: Engine, tau: float = 0.5) -> int:
    """Learned routing of e's resident stream; returns the serialized-lane count."""
    return e.bank_route_learned(tau)


TAUS = (0.002, 0.005, 0.01, 0.02, 0.05, 0.1, 0.2, 0.3, 0.5)


def calibrate(plain: Engine, routed: Engine, thetas: Iterable[float] = (0.6, 0.9, 1.2),
              taus: Iterable[float] = TAUS, n_txns: int = 1 << 17, seed: int = 2000,
              init: int = INIT, make_bank=c1_bank) -> dict:
    """Chooses the routing threshold on the training skews: for each skew, a history
    stream refreshes the counters (plain engine, arrival order), then a validation stream
    is routed at every threshold on the routed engine; the threshold with the highest
    geometric-mean committed txns/s wins.  Returns {"tau", "table"}."""
    thetas, taus = tuple(thetas), tuple(taus)
    logs = {t: [] for t in taus}
    table = []
    for i, th in enumerate(thetas):
        history(plain, make_bank(n_txns, seed + 2 * i, th), init)
        transfer_state(plain, routed)
        routed.bank_generate(make_bank(n_txns, seed + 2 * i + 1, th))
        for t in taus:
            nh = routed.bank_route_learned(t)
            routed.fill(init)
            st = routed.bank_run()
            logs[t].append(np.log(max(st["txn_per_s"], 1.0)))
            table.append({"theta": th, "tau": t, "routed": nh, "txn_per_s": st["txn_per_s"],
                          "abort_rate": st["abort_rate"]})
    best = max(taus, key=lambda t: float(np.mean(logs[t])))
    return {"tau": best, "table": table}
