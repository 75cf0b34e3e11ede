"""CPU restatement of the learned contention predictor (numpy).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, as the checker.

The predictor has no counterpart in the reference (its abort record is the
per-class totals of Statistics, bank/csmv_sh/API.cuh:112-119, and its drivers
dispatch in arrival order, bank/csmv_sh/exe.cu:142-269).  What is restated here
is the engine's own definition:

  abort counters   A[h0(x) mod 4096] += failed attempts of each transfer touching x
  features         f[h1(x) mod 32] += 1 (written item) or 1/2 (read-only item);
                   f[32 + h2(x) mod 31] += floor(log2(1 + A[h0(x)])) / 8;
                   f[63] = max over the items of floor(log2(1 + A[h0(x)])) / 8
  network          z = w2 . relu(W1^T f + b1) + b2, p = sigmoid(z), W1 in bf16
  loss             mean binary cross-entropy of "aborted at least once"
  routing          stable partition of the stream: transfers with p > tau first

How each part is pinned (tests/test_oracle_pred.py): the gradient against
central finite differences of the float64 loss; the forward pass against the
closed form of a network with W1 = 0 and against a hand-evaluated two-unit
case; the features against their closed-form sums (the read/write part sums to
the item weights, the hashed abort part to sum_j lg(A)/8, the last feature to
max_j lg(A)/8); bf16 rounding against
hand-picked ties; routing against a brute-force definition; Adam's first step
against its closed form (each parameter moves by lr against its gradient's sign).
"""
from __future__ import annotations

import numpy as np

N_FEAT = 64
N_HIDDEN = 64
N_PARAMS = N_FEAT * N_HIDDEN + N_HIDDEN + N_HIDDEN + 1
B1 = N_FEAT * N_HIDDEN
W2 = B1 + N_HIDDEN
B2 = W2 + N_HIDDEN
ABORT_BUCKETS = 4096
FEAT_ITEMS = 4
SALT_RW = np.uint32(0xA5A5A5A5)
SALT_AB = np.uint32(0x5A5A5A5A)
KIND_TRANSFER = 0


def feat_mix(x):
    """32-bit xor-shift-multiply finaliser ("lowbias32" constants)."""
    x = np.asarray(x, dtype=np.uint32).copy()
    with np.errstate(over="ignore"):
        x ^= x >> np.uint32(16)
        x *= np.uint32(0x7FEB352D)
        x ^= x >> np.uint32(15)
        x *= np.uint32(0x846CA68B)
        x ^= x >> np.uint32(16)
    return x


def log2p1(a):
    """floor(log2(1 + a)) for uint32 a, exactly (integer bit length)."""
    a = np.asarray(a, dtype=np.uint64) + np.uint64(1)
    out = np.zeros(a.shape, dtype=np.int64)
    for b in range(33):
        out += (a >> np.uint64(b + 1)) > 0
    return out


def bf16_bits(x):
    """float -> bf16 bits, round to nearest even via float32 (finite inputs)."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1))
    return (u >> np.uint64(16)).astype(np.uint16)


def bf16_value(bits):
    return (np.asarray(bits, dtype=np.uint16).astype(np.uint32) << np.uint32(16)).view(np.float32)


def transfer_items(op, addr, T, k):
    o = int(op[k])
    nr = min((o >> 16) & 0xFF, FEAT_ITEMS)
    return [int(addr[j * T + k]) for j in range(nr)]


def abort_fold(op, addr, T, log_ab, counters=None, decay_shift=32):
    """counters >> decay_shift, then += the abort log of the transfers."""
    c = np.zeros(ABORT_BUCKETS, dtype=np.uint64) if counters is None else \
        np.asarray(counters, dtype=np.uint64).copy()
    c = np.zeros_like(c) if decay_shift >= 32 else c >> np.uint64(decay_shift)
    op = np.asarray(op, dtype=np.uint32)
    ab = np.asarray(log_ab, dtype=np.uint64)
    sel = ((op & 0xFF) == KIND_TRANSFER) & (ab > 0)
    nr = np.minimum((op >> 16) & 0xFF, FEAT_ITEMS)
    for j in range(FEAT_ITEMS):
        m = sel & (nr > j)
        if not m.any():
            continue
        x = np.asarray(addr[j * T:(j + 1) * T], dtype=np.uint32)[m]
        np.add.at(c, feat_mix(x) & np.uint32(ABORT_BUCKETS - 1), ab[m])
    return (c & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def features(op, addr, T, idx, counters):
    """Feature rows (float64, exact multiples of 1/8) of transactions idx."""
    op = np.asarray(op, dtype=np.uint32)
    idx = np.asarray(idx, dtype=np.int64)
    X = np.zeros((idx.size, N_FEAT), dtype=np.float64)
    o = op[idx]
    tr = (o & 0xFF) == KIND_TRANSFER
    nr = np.where(tr, np.minimum((o >> 16) & 0xFF, FEAT_ITEMS), 0)
    lg_table = log2p1(np.asarray(counters, dtype=np.uint32))
    rows = np.arange(idx.size)
    lgmax = np.zeros(idx.size, dtype=np.int64)
    for j in range(FEAT_ITEMS):
        m = nr > j
        if not m.any():
            continue
        x = np.asarray(addr, dtype=np.uint32)[j * T + idx[m]]
        w = 1.0 if j < 2 else 0.5
        np.add.at(X, (rows[m], (feat_mix(x ^ SALT_RW) & np.uint32(31)).astype(np.int64)), w)
        lg = lg_table[(feat_mix(x) & np.uint32(ABORT_BUCKETS - 1)).astype(np.int64)]
        np.add.at(X, (rows[m], 32 + (feat_mix(x ^ SALT_AB) % np.uint32(31)).astype(np.int64)),
                  lg * 0.125)
        lgmax[m] = np.maximum(lgmax[m], lg)
    X[:, 63] = np.where(tr, lgmax * 0.125, 0.0)
    return X


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    return (p[:B1].reshape(N_FEAT, N_HIDDEN), p[B1:W2], p[W2:B2], p[B2])


def forward(X, params, round_w1=True):
    """Logits z and probabilities p; W1 rounded to bf16 as the device stores it."""
    W1, b1, w2, b2 = unpack(params)
    if round_w1:
        W1 = bf16_value(bf16_bits(W1.astype(np.float32))).astype(np.float64)
    X = np.asarray(X, dtype=np.float64)
    A = X @ W1 + b1
    H = np.maximum(A, 0.0)
    z = H @ w2 + b2
    return z, 1.0 / (1.0 + np.exp(-z)), A, H


def loss(X, y, params, round_w1=False):
    z, _, _, _ = forward(X, params, round_w1)
    y = np.asarray(y, dtype=np.float64)
    return float(np.mean(np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z)))))


def grad(X, y, params, round_w1=True, round_da=True):
    """Gradient of the mean cross-entropy.  With round_da the hidden-layer gradient dA
    is rounded to bf16 before the weight product, as the device's MFMA operand is."""
    W1, b1, w2, b2 = unpack(params)
    z, p, A, H = forward(X, params, round_w1)
    y = np.asarray(y, dtype=np.float64)
    n = len(y)
    dz = (p - y) / n
    act = A > 0
    dA = np.where(act, dz[:, None] * w2[None, :], 0.0)
    dA_w = bf16_value(bf16_bits(dA.astype(np.float32))).astype(np.float64) if round_da else dA
    g = np.zeros(N_PARAMS, dtype=np.float64)
    g[:B1] = (np.asarray(X, dtype=np.float64).T @ dA_w).reshape(-1)
    g[B1:W2] = dA.sum(axis=0)
    g[W2:B2] = (H * dz[:, None]).sum(axis=0)
    g[B2] = dz.sum()
    L = float(np.mean(np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z)))))
    return g, L


def adam(params, m, v, g, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One bias-corrected Adam step (float64)."""
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    mh = m / (1 - beta1 ** t)
    vh = v / (1 - beta2 ** t)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v


def route(op, p, tau):
    """Stable partition: (normal lane, serialized lane) lists of stream indices."""
    op = np.asarray(op, dtype=np.uint32)
    hot = ((op & 0xFF) == KIND_TRANSFER) & (np.asarray(p) > tau)
    k = np.arange(op.size, dtype=np.uint32)
    return k[~hot], k[hot]
