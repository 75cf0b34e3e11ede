"""GPU parity of a caller-supplied kernel: the external bank driver (tests/driver/
c1_driver.hip, written only against include/) runs the bank-plain configuration through
csmv_kernel_prepare / csmv_kernel_finish with the device API's per-item tx_read /
tx_read_only, and its log replays serially in commit order (oracle)."""
import numpy as np
import pytest

import c1drv
import csmv
import orc

pytestmark = pytest.mark.gpu

INIT = 1000


@pytest.mark.parametrize("variant,n_txns", [("stamp", 1 << 20), ("vanilla", 1 << 16),
                                            ("batchedAtrInsert", 1 << 16),
                                            ("collabValidation", 1 << 16)])
def test_external_driver_bank_plain(variant, n_txns):
    """The bank-plain configuration (4096 accounts, 2-account transfers, 32-account audits,
    90/10, 64 client waves x 64 lanes + the server workgroup, depth 4) driven by the
    external kernel: every transaction commits and the log replays exactly."""
    bank = csmv.Bank(n_txns=n_txns, seed=17, n_reads=2, n_reads_ro=32, pct_ro=10)
    print(f"\n[case] external driver {variant} {n_txns} txns", flush=True)
    with csmv.Engine(4096, depth=4, atr_entries=4096, client_waves=64, wg_waves=6,
                     variant=variant, timeout_ms=20000) as e:
        e.bank_generate(bank)
        op, addr = e.bank_read_stream()
        e.fill(INIT)
        st, ts, v, ab = c1drv.run(e, op, addr, bank.rmax)
        final = e.read_state()
    spec = orc.bank_spec(4096, n_txns, seed=bank.seed, n_reads=2, n_reads_ro=32, pct_ro=10)
    rep, _ = orc.bank_replay(spec, op, addr, INIT, ts, v, final)
    print("[stats]", {k: st[k] for k in ("commits", "aborts", "abort_rate", "kernel_ms",
                                         "txn_per_s")}, flush=True)
    assert rep["fails"] == 0, rep
    assert st["commits"] == n_txns and st["error"] == 0
    assert st["commits_update"] == rep["n_updates"] == st["final_ts"]
    assert ab.sum() == st["aborts"]


def test_prepare_rejects_wrong_buffer_sizes():
    """csmv_kernel_prepare checks the caller's context / layout buffer sizes and the
    order of calls before it touches the device."""
    import ctypes as C
    print("\n[case] kernel_prepare argument checks", flush=True)
    L = csmv.lib()
    L.csmv_kernel_prepare.argtypes = [C.c_void_p] * 2 + [C.c_void_p, C.c_size_t, C.c_void_p,
                                                          C.c_size_t, C.c_void_p]
    L.csmv_kernel_finish.argtypes = [C.c_void_p, C.c_void_p]

    class Spec(C.Structure):
        _fields_ = [("kernel", C.c_void_p), ("reads", C.c_uint32), ("writes", C.c_uint32)]

    buf = (C.c_uint8 * 4096)()
    lnch = (C.c_uint8 * 64)()
    with csmv.Engine(4096, depth=4, client_waves=8) as e:
        e.fill(INIT)
        sp = Spec(C.cast(buf, C.c_void_p).value, 2, 2)
        assert L.csmv_kernel_prepare(e._h, C.byref(sp), buf, 7, buf, 7, lnch) == 2
        assert L.csmv_kernel_finish(e._h, None) == 6
