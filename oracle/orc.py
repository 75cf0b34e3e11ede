"""ctypes wrapper of the CPU oracle (oracle/liborcl.so).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborcl.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


class BankSpec(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("n_txns", C.c_uint64), ("n_accounts", C.c_uint32),
                ("n_reads", C.c_uint32), ("n_reads_ro", C.c_uint32), ("pct_ro", C.c_uint32),
                ("dist", C.c_uint32), ("amount_max", C.c_uint32), ("n_reads_min", C.c_uint32),
                ("theta", C.c_double)]


class ReplayReport(C.Structure):
    _fields_ = [("n_updates", C.c_uint64), ("n_audits", C.c_uint64), ("bad_ts", C.c_uint64),
                ("bad_snapshot", C.c_uint64), ("bad_reads", C.c_uint64),
                ("bad_audits", C.c_uint64), ("bad_final", C.c_uint64), ("total", C.c_int64),
                ("expected_total", C.c_int64), ("first_bad_txn", C.c_int64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class CpuCfg(C.Structure):
    _fields_ = [("depth", C.c_uint32), ("atr_entries", C.c_uint32), ("n_clients", C.c_uint32),
                ("pin", C.c_uint32)]


class CpuStats(C.Structure):
    _fields_ = [("commits_update", C.c_uint64), ("commits_ro", C.c_uint64),
                ("aborts_dataage", C.c_uint64), ("aborts_readwrite", C.c_uint64),
                ("aborts_recordage", C.c_uint64), ("seconds", C.c_double),
                ("final_ts", C.c_uint32), ("threads", C.c_uint32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.orc_splitmix64.argtypes = [C.c_uint64]
        L.orc_splitmix64.restype = C.c_uint64
        L.orc_rand_r_fnc.argtypes = [C.POINTER(C.c_uint64)]
        L.orc_rand_r_fnc.restype = C.c_uint32
        L.orc_txn_state.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_txn_state.restype = C.c_uint64
        L.orc_zipf_cdf.argtypes = [C.c_uint32, C.c_double, P]
        L.orc_zipf_rank.argtypes = [P, C.c_uint32, C.c_uint32]
        L.orc_zipf_rank.restype = C.c_uint32
        L.orc_scramble.argtypes = [C.c_uint32, C.c_uint32]
        L.orc_scramble.restype = C.c_uint32
        L.orc_bank_gen.argtypes = [C.POINTER(BankSpec), P, P, P]
        L.orc_bank_replay.argtypes = [C.POINTER(BankSpec), P, P, C.c_int32, P, P, P, P,
                                      C.POINTER(ReplayReport)]
        L.orc_bank_replay.restype = C.c_int64
        L.orc_bank_serial.argtypes = [C.POINTER(BankSpec), P, P, C.c_int32, P, P, P, P]
        L.orc_bank_cpu_csmv.argtypes = [C.POINTER(BankSpec), P, P, C.c_int32, C.c_uint64,
                                        C.POINTER(CpuCfg), P, P, P, C.POINTER(CpuStats)]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


def splitmix64(x: int) -> int:
    return lib().orc_splitmix64(x)


def rand_r_fnc(state: int):
    s = C.c_uint64(state)
    r = lib().orc_rand_r_fnc(C.byref(s))
    return r, s.value


def zipf_cdf(n: int, theta: float) -> np.ndarray:
    out = np.empty(n, dtype=np.uint32)
    assert lib().orc_zipf_cdf(n, theta, _p(out)) == 0
    return out


def zipf_rank(cdf: np.ndarray, r: int) -> int:
    return lib().orc_zipf_rank(_p(cdf), cdf.size, r)


def scramble(rank: int, n: int) -> int:
    return lib().orc_scramble(rank, n)


def bank_spec(n_accounts, n_txns, seed=1, n_reads=2, n_reads_ro=32, pct_ro=10, zipf=False,
              theta=0.99, amount_max=10, n_reads_min=0) -> BankSpec:
    return BankSpec(seed, n_txns, n_accounts, n_reads, n_reads_ro, pct_ro, 1 if zipf else 0,
                    amount_max, n_reads_min, theta)


def rmax(spec: BankSpec) -> int:
    return max(spec.n_reads, spec.n_reads_ro)


def bank_gen(spec: BankSpec):
    cdf = zipf_cdf(spec.n_accounts, spec.theta) if spec.dist == 1 else None
    op = np.empty(spec.n_txns, dtype=np.uint32)
    addr = np.empty(rmax(spec) * spec.n_txns, dtype=np.uint32)
    assert lib().orc_bank_gen(C.byref(spec), _p(cdf), _p(op), _p(addr)) == 0
    return op, addr


def bank_replay(spec, op, addr, init, log_ts, log_v, final=None):
    rep = ReplayReport()
    replayed = np.empty(spec.n_accounts, dtype=np.int32)
    fails = lib().orc_bank_replay(C.byref(spec), _p(op), _p(addr), int(init), _p(log_ts),
                                  _p(log_v), _p(final), _p(replayed), C.byref(rep))
    d = rep.as_dict()
    d["fails"] = fails
    return d, replayed


def bank_serial(spec, op, addr, init, order=None):
    state = np.empty(spec.n_accounts, dtype=np.int32)
    log_ts = np.zeros(spec.n_txns, dtype=np.uint32)
    log_v = np.zeros(spec.n_reads * spec.n_txns, dtype=np.int32)
    o = None if order is None else np.ascontiguousarray(order, dtype=np.uint64)
    assert lib().orc_bank_serial(C.byref(spec), _p(op), _p(addr), int(init), _p(o), _p(state),
                                 _p(log_ts), _p(log_v)) == 0
    return state, log_ts, log_v


def bank_cpu_csmv(spec, op, addr, init, n_run=None, depth=4, atr_entries=4096, n_clients=4,
                  pin=False):
    n_run = spec.n_txns if n_run is None else n_run
    cfg = CpuCfg(depth, atr_entries, n_clients, 1 if pin else 0)
    st = CpuStats()
    state = np.empty(spec.n_accounts, dtype=np.int32)
    log_ts = np.zeros(spec.n_txns, dtype=np.uint32)
    log_v = np.zeros(spec.n_reads * spec.n_txns, dtype=np.int32)
    rc = lib().orc_bank_cpu_csmv(C.byref(spec), _p(op), _p(addr), int(init), n_run,
                                 C.byref(cfg), _p(state), _p(log_ts), _p(log_v), C.byref(st))
    assert rc == 0, rc
    return state, log_ts, log_v, st.as_dict()


class MemcdSpec(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("n_reqs", C.c_uint64), ("n_keys", C.c_uint64),
                ("n_sets", C.c_uint32), ("n_ways", C.c_uint32), ("pct_get", C.c_uint32),
                ("batch", C.c_uint32), ("dist", C.c_uint32), ("reserved", C.c_uint32),
                ("theta", C.c_double)]


class MemcdReport(C.Structure):
    _fields_ = [(f, C.c_uint64) for f in ("n_updates", "n_readonly", "hits", "misses", "sets",
                                          "bad_ts", "bad_snapshot", "bad_result", "bad_kind",
                                          "bad_final")] + [("first_bad_req", C.c_int64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def _memcd_lib():
    L = lib()
    if not getattr(L, "_memcd_ready", False):
        P = C.c_void_p
        L.orc_memcd_key.argtypes = [C.c_uint32]
        L.orc_memcd_key.restype = C.c_uint64
        L.orc_memcd_set_of.argtypes = [C.c_uint64, C.c_uint32]
        L.orc_memcd_set_of.restype = C.c_uint32
        L.orc_memcd_value_word.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32]
        L.orc_memcd_value_word.restype = C.c_uint32
        L.orc_memcd_gen.argtypes = [C.POINTER(MemcdSpec), P, P, P]
        L.orc_memcd_serial.argtypes = [C.POINTER(MemcdSpec), P, P, P, P, P, P, P]
        L.orc_memcd_replay.argtypes = [C.POINTER(MemcdSpec), P, P, P, P, P, P,
                                       C.POINTER(MemcdReport)]
        L.orc_memcd_replay.restype = C.c_int64
        L.orc_memcd_serial_from.argtypes = [C.POINTER(MemcdSpec), P, P, P, P, P, P, P, P]
        L.orc_memcd_replay_from.argtypes = [C.POINTER(MemcdSpec), P, P, P, P, P, P, P,
                                            C.POINTER(MemcdReport)]
        L.orc_memcd_replay_from.restype = C.c_int64
        L._memcd_ready = True
    return L


def memcd_spec(n_reqs, seed=1, n_keys=10_000_000, n_sets=1 << 17, n_ways=8, pct_get=95,
               batch=65536, zipf=True, theta=0.99) -> MemcdSpec:
    return MemcdSpec(seed, n_reqs, n_keys, n_sets, n_ways, pct_get, batch, 1 if zipf else 0, 0,
                     theta)


def memcd_gen(spec: MemcdSpec):
    L = _memcd_lib()
    cdf = zipf_cdf(spec.n_keys, spec.theta) if spec.dist == 1 else None
    key = np.empty(spec.n_reqs, dtype=np.uint64)
    op = np.empty(spec.n_reqs, dtype=np.uint32)
    assert L.orc_memcd_gen(C.byref(spec), _p(cdf), _p(key), _p(op)) == 0
    return key, op


def memcd_serial(spec, key, op, order=None, init=None):
    L = _memcd_lib()
    items = np.empty(12 * spec.n_sets * spec.n_ways, dtype=np.int32)
    ts = np.zeros(spec.n_reqs, dtype=np.uint32)
    res = np.zeros(spec.n_reqs, dtype=np.uint32)
    vsum = np.zeros(spec.n_reqs, dtype=np.uint32)
    o = None if order is None else np.ascontiguousarray(order, dtype=np.uint64)
    i0 = None if init is None else np.ascontiguousarray(init, dtype=np.int32)
    assert L.orc_memcd_serial_from(C.byref(spec), _p(i0), _p(key), _p(op), _p(o), _p(items),
                                   _p(ts), _p(res), _p(vsum)) == 0
    return items, ts, res, vsum


def memcd_replay(spec, key, op, ts, res, vsum, final=None, init=None):
    L = _memcd_lib()
    rep = MemcdReport()
    i0 = None if init is None else np.ascontiguousarray(init, dtype=np.int32)
    fails = L.orc_memcd_replay_from(C.byref(spec), _p(i0), _p(key), _p(op), _p(ts), _p(res),
                                    _p(vsum), _p(final), C.byref(rep))
    d = rep.as_dict()
    d["fails"] = fails
    return d
