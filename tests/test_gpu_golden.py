"""The device against the committed golden fixtures: stream generators bit-exact, the
predictor's forward probabilities within 1e-4 and gradients within 1e-2 of the largest
(the fixture is the exact float64 gradient; the device rounds dA to bf16 for the weight
product, a relative 2^-9 per element, and accumulates in float32)."""
import json
import os

import numpy as np
import pytest

import csmv

pytestmark = pytest.mark.gpu

FX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fixtures.json")))


@pytest.mark.parametrize("name", ["bank_plain", "bank_scale", "bank_serving"])
def test_device_bank_stream_prefix(name):
    f = FX[name]
    sp = f["spec"]
    print(f"\n[case] golden {name}", flush=True)
    # the fixture spec lists what differs from the stream defaults (32-read audits, 10 %)
    b = csmv.Bank(n_txns=sp["n_txns"], seed=sp["seed"], n_reads=sp["n_reads"],
                  n_reads_ro=sp.get("n_reads_ro", 32), pct_ro=sp.get("pct_ro", 10),
                  zipf=sp.get("zipf", False),
                  theta=sp.get("theta", 0.99), n_reads_min=sp.get("n_reads_min", 0))
    with csmv.Engine(sp["n_accounts"], depth=4, client_waves=8) as e:
        e.bank_generate(b)
        op, addr = e.bank_read_stream()
    assert op.tolist() == f["op"] and addr.tolist() == f["addr"]


def test_device_memcd_stream_prefix():
    f = FX["memcd_config"]
    print("\n[case] golden memcd stream", flush=True)
    m = csmv.Memcd(n_reqs=f["n_reqs"], seed=f["seed"])
    with csmv.Engine(m.n_items, depth=2, client_waves=8) as e:
        e.memcd_generate(m)
        key, op = e.memcd_read_stream()
    assert [int(k) for k in key] == f["key"] and op.tolist() == f["op"]


def test_device_mlp_tiny():
    f = FX["mlp_tiny"]
    print("\n[case] golden tiny MLP forward/backward", flush=True)
    with csmv.Engine(4096, depth=4, client_waves=8) as e:
        e.pred_set_params(np.array(f["params"], dtype=np.float32))
        X = np.array(f["X_bf16"], dtype=np.uint16)
        p = e.pred_forward(X)
        assert np.max(np.abs(p - np.array(f["p"]))) < 1e-4
        g, loss = e.pred_grad(X, np.array(f["y"], dtype=np.float32))
    want = np.array(f["grad"])
    assert np.max(np.abs(g - want)) <= 1e-2 * np.max(np.abs(want))
    assert abs(loss - f["loss"]) <= 1e-5 * abs(f["loss"])
