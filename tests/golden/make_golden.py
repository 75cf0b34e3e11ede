"""Writes tests/golden/fixtures.json: small input/output vectors for the parity tests.

Sources, per fixture:
  rand_r          glibc rand_r (libc), the generator RAND_R_FNC reduces to
  splitmix64      the published SplitMix64 sequence seeded with 0
  zipf_cdf        floor(2^31 * H_theta(i+1) / H_theta(n)) with math.fsum
  bank_* / memcd  the first transactions of the configured streams, from the oracle
                  (regression vectors: the device generators must reproduce them)
  mlp_tiny        a 3-row network evaluated in float64; the gradient is accepted only
                  after it matches central finite differences

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import mlp_oracle as M  # noqa: E402
import orc  # noqa: E402


def rand_r_vectors():
    libc = ctypes.CDLL("libc.so.6")
    libc.rand_r.argtypes = [ctypes.POINTER(ctypes.c_uint)]
    libc.rand_r.restype = ctypes.c_int
    out = []
    for s in (0, 1, 7, 12345, 2**31 - 1, 2**32 - 1):
        st = ctypes.c_uint(s)
        vals = [libc.rand_r(ctypes.byref(st)) for _ in range(4)]
        out.append({"state": s, "values": vals})
    return out


def zipf_points(n, theta, idx):
    terms = [(i + 1) ** -theta for i in range(n)]
    total = math.fsum(terms)
    return [{"i": i, "cdf": int(math.floor(2**31 * math.fsum(terms[:i + 1]) / total))}
            for i in idx]


def stream(n_accounts, k, **kw):
    spec = orc.bank_spec(n_accounts, k, **kw)
    op, addr = orc.bank_gen(spec)
    return {"spec": {"n_accounts": n_accounts, "n_txns": k, **kw}, "op": op.tolist(),
            "addr": addr.tolist()}


def mlp_tiny():
    rng = np.random.default_rng(2026)
    p = rng.uniform(-0.4, 0.4, M.N_PARAMS).astype(np.float32)
    p[:M.B1] = M.bf16_value(M.bf16_bits(p[:M.B1]))
    X = (rng.integers(0, 24, (3, 64)) * (rng.random((3, 64)) < 0.3)) / 8
    y = np.array([1.0, 0.0, 1.0])
    p64 = p.astype(np.float64)
    z, prob, _, _ = M.forward(X, p64)
    g, L = M.grad(X, y, p64, round_w1=False, round_da=False)
    h = 1e-6
    for i in list(range(0, M.B1, 97)) + list(range(M.B1, M.N_PARAMS)):
        a, b = p64.copy(), p64.copy()
        a[i] += h
        b[i] -= h
        fd = (M.loss(X, y, a) - M.loss(X, y, b)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-8 + 1e-5 * abs(g[i]), (i, fd, g[i])
    return {"params": p.tolist(), "X_bf16": M.bf16_bits(X).astype(int).tolist(), "y": y.tolist(),
            "p": prob.tolist(), "z": z.tolist(), "loss": L, "grad": g.tolist()}


def main():
    fx = {
        "rand_r": rand_r_vectors(),
        "splitmix64": {"inputs": [(i * 0x9E3779B97F4A7C15) & (2**64 - 1) for i in range(3)],
                       "outputs": [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]},
        "zipf_cdf": {"n": 4096, "theta": 0.99,
                     "points": zipf_points(4096, 0.99, [0, 1, 9, 99, 1023, 4095])},
        "bank_plain": stream(4096, 32, seed=1, n_reads=2, n_reads_ro=32, pct_ro=10),
        "bank_scale": stream(1 << 20, 32, seed=1, n_reads=4, n_reads_ro=4, pct_ro=10, zipf=True,
                             theta=0.99),
        "bank_serving": stream(1 << 20, 16, seed=7, n_reads=64, n_reads_min=2, pct_ro=0,
                               zipf=True, theta=0.99),
        "mlp_tiny": mlp_tiny(),
    }
    spec = orc.memcd_spec(32, seed=1)
    key, op = orc.memcd_gen(spec)
    fx["memcd_config"] = {"n_reqs": 32, "seed": 1, "key": [int(x) for x in key],
                          "op": op.tolist()}
    with open(os.path.join(HERE, "fixtures.json"), "w") as f:
        json.dump(fx, f, separators=(",", ":"))
    print("wrote", os.path.join(HERE, "fixtures.json"))


if __name__ == "__main__":
    main()
