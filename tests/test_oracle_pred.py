"""Pins the predictor oracle (oracle/mlp_oracle.py) by mathematics, not by its own formulas:
finite differences, closed forms, hand-evaluated cases and brute force."""
import math

import numpy as np

import mlp_oracle as M
import orc


def test_bf16_round_to_nearest_even():
    one = 0x3F80
    cases = [(1.0, one), (-1.0, 0xBF80), (0.0, 0),
             (1.0 + 2.0 ** -8, one),                  # tie -> even (1.0)
             (1.0 + 3 * 2.0 ** -8, one + 2),          # tie -> even (1 + 2^-6)
             (1.0 + 2.0 ** -8 + 2.0 ** -20, one + 1),  # above the tie
             (1.0 + 2.0 ** -9, one),                  # below the tie
             (0.125, 0x3E00), (15.5, 0x4178)]
    for x, bits in cases:
        assert int(M.bf16_bits(np.float32(x))) == bits, (x, hex(int(M.bf16_bits(np.float32(x)))))
    # every multiple of 1/8 below 32 is exact in bf16 (the features' range)
    v = np.arange(0, 256, dtype=np.float32) / 8
    assert np.array_equal(M.bf16_value(M.bf16_bits(v)), v)


def test_log2p1_is_bit_length():
    a = np.array([0, 1, 2, 3, 4, 7, 8, 1000, 2 ** 31, 2 ** 32 - 1], dtype=np.uint32)
    expect = [(int(x) + 1).bit_length() - 1 for x in a]
    assert list(M.log2p1(a)) == expect
    assert expect[-1] == 32


def test_feat_mix_is_a_bijection_on_a_range():
    x = np.arange(1 << 18, dtype=np.uint32)
    h = M.feat_mix(x)
    assert np.unique(h).size == x.size
    assert int(M.feat_mix(np.uint32(0))) == 0   # the finaliser fixes zero


def _stream(n_txns, theta=0.9, seed=7, n_reads=2, n_reads_ro=32):
    spec = orc.bank_spec(4096, n_txns, seed=seed, n_reads=n_reads, n_reads_ro=n_reads_ro,
                         zipf=True, theta=theta)
    op, addr = orc.bank_gen(spec)
    return spec, op, addr


def test_features_closed_form_sums():
    rng = np.random.default_rng(1)
    for n_reads in (2, 4):
        spec, op, addr = _stream(3000, n_reads=n_reads)
        T = spec.n_txns
        counters = rng.integers(0, 1 << 20, M.ABORT_BUCKETS).astype(np.uint32)
        counters[:100] = 0
        X = M.features(op, addr, T, np.arange(T), counters)
        for k in range(0, T, 7):
            o = int(op[k])
            if o & 0xFF:
                assert not X[k].any(), "audits have no features"
                continue
            items = [int(addr[j * T + k]) for j in range(n_reads)]
            # read/write part: the item weights (1 for the two written, 1/2 for the others)
            assert X[k, :32].sum() == 2 + 0.5 * (n_reads - 2)
            # abort part: the hashed terms sum floor(log2(1 + A)) / 8 over the items, the
            # last feature is its largest term
            lgs = []
            for x in items:
                a = int(counters[int(M.feat_mix(np.uint32(x))) & 4095])
                lgs.append((a + 1).bit_length() - 1)
            assert X[k, 32:63].sum() == sum(lgs) / 8
            assert X[k, 63] == max(lgs) / 8
            for x in items:
                assert X[k, 32 + int(M.feat_mix(np.uint32(x) ^ M.SALT_AB)) % 31] > 0 or \
                    (int(counters[int(M.feat_mix(np.uint32(x))) & 4095]) == 0)
            for x in items:
                assert X[k, int(M.feat_mix(np.uint32(x) ^ M.SALT_RW)) & 31] >= 0.5
        # exact in bf16
        assert np.array_equal(M.bf16_value(M.bf16_bits(X)).astype(np.float64), X)


def test_features_zero_counters_have_no_abort_part():
    spec, op, addr = _stream(500)
    X = M.features(op, addr, spec.n_txns, np.arange(spec.n_txns),
                   np.zeros(M.ABORT_BUCKETS, dtype=np.uint32))
    assert not X[:, 32:].any()
    ones = np.full(M.ABORT_BUCKETS, 3, dtype=np.uint32)   # lg(4) = 2 -> 1/4 per item
    X = M.features(op, addr, spec.n_txns, np.arange(spec.n_txns), ones)
    tr = (op & 0xFF) == 0
    assert np.all(X[tr, 32:63].sum(axis=1) == 0.5)
    assert np.all(X[tr, 63] == 0.25) and not X[~tr].any()


def test_abort_fold_brute_force():
    spec, op, addr = _stream(2000, n_reads=4, n_reads_ro=4)
    T = spec.n_txns
    rng = np.random.default_rng(2)
    ab = rng.integers(0, 4, T).astype(np.uint32)
    prev = rng.integers(0, 1000, M.ABORT_BUCKETS).astype(np.uint32)
    got = M.abort_fold(op, addr, T, ab, prev, decay_shift=1)
    want = [int(c) >> 1 for c in prev]
    for k in range(T):
        if int(op[k]) & 0xFF or ab[k] == 0:
            continue
        for j in range(4):
            want[int(M.feat_mix(np.uint32(addr[j * T + k]))) & 4095] += int(ab[k])
    assert list(got) == want
    assert not M.abort_fold(op, addr, T, np.zeros(T, np.uint32), prev, decay_shift=32).any()


def _params(rng, scale=0.3):
    p = rng.uniform(-scale, scale, M.N_PARAMS)
    return p


def test_forward_closed_forms():
    rng = np.random.default_rng(3)
    p = _params(rng)
    p[:M.B1] = 0.0                                  # W1 = 0: every row gives the same logit
    X = rng.integers(0, 32, (10, 64)) / 8
    z, prob, _, _ = M.forward(X, p)
    want = np.maximum(p[M.B1:M.W2], 0) @ p[M.W2:M.B2] + p[M.B2]
    assert np.allclose(z, want, rtol=0, atol=1e-12)
    assert np.allclose(prob, 1 / (1 + math.exp(-want)))
    # hand-evaluated: x = e_0 + 2 e_1, units 0 and 1 only
    q = np.zeros(M.N_PARAMS)
    q[0 * 64 + 0] = 1.0      # W1[0][0]
    q[1 * 64 + 0] = -0.5     # W1[1][0]
    q[1 * 64 + 1] = 2.0      # W1[1][1]
    q[M.B1 + 0] = 0.25
    q[M.B1 + 1] = -1.0
    q[M.W2 + 0] = 3.0
    q[M.W2 + 1] = -2.0
    q[M.B2] = 0.5
    x = np.zeros((1, 64))
    x[0, 0], x[0, 1] = 1.0, 2.0
    # a0 = 1 - 1 + 0.25 = 0.25 ; a1 = 4 - 1 = 3 ; z = 3*0.25 - 2*3 + 0.5 = -4.75
    z, _, _, _ = M.forward(x, q)
    assert z[0] == -4.75


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(4)
    n = 40
    X = rng.integers(0, 24, (n, 64)) / 8
    y = (rng.random(n) < 0.4).astype(np.float64)
    p = _params(rng, 0.2)
    g, L = M.grad(X, y, p, round_w1=False, round_da=False)
    assert abs(L - M.loss(X, y, p)) < 1e-12
    idx = np.concatenate([rng.choice(M.B1, 150, replace=False), np.arange(M.B1, M.N_PARAMS)])
    h = 1e-6
    for i in idx:
        a, b = p.copy(), p.copy()
        a[i] += h
        b[i] -= h
        fd = (M.loss(X, y, a) - M.loss(X, y, b)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(g[i]), (i, fd, g[i])


def test_rounded_gradient_stays_close_to_exact():
    rng = np.random.default_rng(5)
    X = rng.integers(0, 24, (300, 64)) / 8
    y = (rng.random(300) < 0.5).astype(np.float64)
    p = _params(rng, 0.2)
    ge, _ = M.grad(X, y, p, round_w1=False, round_da=False)
    gr, _ = M.grad(X, y, p, round_w1=False, round_da=True)
    assert np.max(np.abs(ge - gr)) <= 1e-2 * np.max(np.abs(ge))
    assert np.array_equal(ge[M.B1:], gr[M.B1:]), "only the W1 product uses the rounded dA"


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(6)
    p = _params(rng)
    g = rng.choice([-1.0, 1.0], M.N_PARAMS) * rng.uniform(0.1, 3.0, M.N_PARAMS)
    lr = 0.01
    new, m, v = M.adam(p, np.zeros_like(p), np.zeros_like(p), g, 1, lr)
    # m_hat = g, v_hat = g^2: each parameter moves by lr against its gradient's sign
    assert np.allclose(new, p - lr * np.sign(g), rtol=0, atol=1e-8)   # eps / |g| <= 1e-7


def test_route_brute_force():
    rng = np.random.default_rng(7)
    op = rng.integers(0, 2, 1000).astype(np.uint32)
    p = rng.random(1000)
    norm, hot = M.route(op, p, 0.6)
    h = [k for k in range(1000) if op[k] == 0 and p[k] > 0.6]
    assert list(hot) == h
    assert list(norm) == [k for k in range(1000) if k not in set(h)]
