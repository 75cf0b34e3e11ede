"""The allocation-free bank run (csmv_bank_run_noalloc) against the engine's csmv_bank_run:
the same stream on caller-owned device buffers, two calls in a row (the second smaller, so
the fixed scratch array is reused with stale contents from the first).

Transfers commute, so the final balances must be bit-identical between the two entry
points whatever the commit order; each entry point's log must also replay serially in its
own commit order (oracle), and the transfer and audit counts must agree."""
import numpy as np
import pytest
import torch

import csmv
import orc

pytestmark = pytest.mark.gpu

INIT = 1000


def _on_device(a: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int32)).to("cuda")


@pytest.mark.parametrize("escalate", [0, 2])
def test_noalloc_matches_bank_run(escalate):
    kw = dict(depth=4, atr_entries=4096, client_waves=16, wg_waves=6, variant="stamp",
              timeout_ms=20000, escalate_after=escalate)
    for n_items, n_txns, seed in ((4096, 1 << 16, 81), (1024, 1 << 13, 82)):
        bank = csmv.Bank(n_txns=n_txns, seed=seed, n_reads=2, n_reads_ro=32, pct_ro=10)
        print(f"\n[case] noalloc vs bank_run: {n_items} accounts, {n_txns} txns, "
              f"escalate_after={escalate}", flush=True)
        with csmv.Engine(n_items, **kw) as e:
            e.fill(INIT)
            e.bank_generate(bank)
            st_e = e.bank_run()
            op, addr = e.bank_read_stream()
            ts_e, v_e, _ = e.bank_read_log()
            fin_e = e.read_state()

        d_init = torch.full((n_items,), INIT, dtype=torch.int32, device="cuda")
        d_final = torch.zeros(n_items, dtype=torch.int32, device="cuda")
        d_op, d_addr = _on_device(op), _on_device(addr)
        d_ts = torch.zeros(n_txns, dtype=torch.int32, device="cuda")
        d_v = torch.zeros(bank.n_reads * n_txns, dtype=torch.int32, device="cuda")
        d_ab = torch.zeros(n_txns, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        bufs = csmv.BankBuffers(d_init.data_ptr(), d_final.data_ptr(), d_op.data_ptr(),
                                d_addr.data_ptr(), n_txns, d_ts.data_ptr(), d_v.data_ptr(),
                                d_ab.data_ptr())
        st_n = csmv.bank_run_noalloc(csmv.config(n_items, **kw), bufs,
                                     torch.cuda.current_stream().cuda_stream)
        fin_n = d_final.cpu().numpy()
        ts_n = d_ts.cpu().numpy().view(np.uint32)
        v_n = d_v.cpu().numpy()
        ab_n = d_ab.cpu().numpy().view(np.uint32)
        print("[stats] bank_run", {k: st_e[k] for k in ("commits", "aborts", "kernel_ms")},
              "noalloc", {k: st_n[k] for k in ("commits", "aborts", "kernel_ms", "escalated")},
              flush=True)

        assert st_n["error"] == 0
        assert st_n["commits"] == st_e["commits"] == n_txns
        assert st_n["commits_update"] == st_e["commits_update"]
        assert st_n["commits_ro"] == st_e["commits_ro"]
        assert np.array_equal(fin_n, fin_e)
        assert int(ab_n.sum()) == st_n["aborts"]
        spec = orc.bank_spec(n_items, n_txns, seed=seed, n_reads=2, n_reads_ro=32, pct_ro=10)
        for ts, v, fin in ((ts_e, v_e, fin_e), (ts_n, v_n, fin_n)):
            rep, _ = orc.bank_replay(spec, op, addr, INIT, ts, v, fin)
            assert rep["fails"] == 0, rep
            assert rep["total"] == n_items * INIT
