"""The oracle against the committed golden fixtures (tests/golden/fixtures.json)."""
import json
import os

import numpy as np

import mlp_oracle as M
import orc

FX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fixtures.json")))


def test_rand_r_vectors():
    for case in FX["rand_r"]:
        s = case["state"]
        for want in case["values"]:
            got, s = orc.rand_r_fnc(s)
            assert got == want


def test_splitmix64_vectors():
    sm = FX["splitmix64"]
    assert [orc.splitmix64(x) for x in sm["inputs"]] == sm["outputs"]


def test_zipf_cdf_points():
    z = FX["zipf_cdf"]
    cdf = orc.zipf_cdf(z["n"], z["theta"])
    for pt in z["points"]:
        assert abs(int(cdf[pt["i"]]) - pt["cdf"]) <= 1


def _stream_matches(name):
    f = FX[name]
    spec = f["spec"]
    full = orc.bank_spec(spec["n_accounts"], spec["n_txns"],
                         **{k: v for k, v in spec.items() if k not in ("n_accounts", "n_txns")})
    op, addr = orc.bank_gen(full)
    assert op.tolist() == f["op"] and addr.tolist() == f["addr"]


def test_bank_streams():
    for name in ("bank_plain", "bank_scale", "bank_serving"):
        _stream_matches(name)


def test_memcd_stream():
    f = FX["memcd_config"]
    key, op = orc.memcd_gen(orc.memcd_spec(f["n_reqs"], seed=f["seed"]))
    assert [int(k) for k in key] == f["key"] and op.tolist() == f["op"]


def test_mlp_tiny():
    f = FX["mlp_tiny"]
    p = np.array(f["params"], dtype=np.float64)
    X = M.bf16_value(np.array(f["X_bf16"], dtype=np.uint16)).astype(np.float64)
    y = np.array(f["y"])
    z, prob, _, _ = M.forward(X, p)
    assert np.allclose(prob, f["p"], rtol=0, atol=1e-12)
    g, L = M.grad(X, y, p, round_w1=False, round_da=False)
    assert np.allclose(g, f["grad"], rtol=0, atol=1e-12) and abs(L - f["loss"]) < 1e-12
