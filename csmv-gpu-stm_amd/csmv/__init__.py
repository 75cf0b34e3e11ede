"""ctypes harness over the gfx950 CSMV engine's C-ABI (include/csmv.h).

The engine is the native library ``libcsmv.so`` built in-tree from
``csmv-gpu-stm_amd/csrc``.  This module only marshals arguments; every
transaction runs in the HIP kernels.  There is no CPU fallback: if the library
is missing or no gfx950 device is visible, the calls raise.

Reference interfaces mirrored here (file:line in DMRNunes/csmv-gpu-stm):
  Engine(...)               TXInit + create_gbc   bank/csmv_sh/API.cu:293-333, sync_lib/one_phase_server.h:13-42
  Engine.fill / load        TXInit data upload    bank/csmv_sh/API.cu:310-322
  Engine.read_state         TXEnd read-back       bank/csmv_sh/API.cu:337-348
  Engine.bank_run           test_fine_grain_offloading   bank/csmv_sh/exe.cu:340-529
  Engine.close              TXEnd frees + free_gbc       bank/csmv_sh/API.cu:354-358
  Engine.abort_fold ...     per-bucket abort log  (the reference keeps per-class totals only,
                            Statistics, bank/csmv_sh/API.cuh:112-119)
  Engine.pred_* / bank_route_learned   learned contention predictor: no reference counterpart
  serve(cache, bank, ...)   persistent serving loop; the reference serves cache batches from a
                            host loop against a persistent server, memcachedGPU/memcdSH.cu:603-651
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libcsmv.so")

# variant bits (csmv.h)
VAL_SCAN, VAL_COLLAB, VAL_STAMP = 0x0, 0x1, 0x2
INSERT_LANE, INSERT_BATCH = 0x0, 0x4
WB_CLIENT, WB_SERVER = 0x0, 0x8
VARIANTS = {
    "vanilla": VAL_SCAN | INSERT_LANE | WB_SERVER,
    "batchedAtrInsert": VAL_SCAN | INSERT_BATCH | WB_CLIENT,
    "clientWriteback": VAL_SCAN | INSERT_BATCH | WB_SERVER,
    "collabValidation": VAL_COLLAB | INSERT_BATCH | WB_CLIENT,
    "stamp": VAL_STAMP | INSERT_BATCH | WB_CLIENT,
}
# the pipelined serialized lane (bank driver, every transfer routed to it): clients stage
# their requests' producer chains, the committer orders each item's request-wide delta with
# one atomic and installs the versions itself
SERIAL_PIPE = 0x10
PRESETS = dict(VARIANTS, stamp_pipe=VARIANTS["stamp"] | SERIAL_PIPE)

STATUS = {0: "ok", 1: "HIP runtime error", 2: "invalid argument", 3: "capacity exceeded",
          4: "device wait timed out", 5: "no gfx950 device", 6: "call out of order",
          7: "transaction attempt limit exceeded", 8: "log file I/O failed or file malformed"}


PRED_PARAMS = 4225      # W1[64][64] f-major, b1[64], w2[64], b2
PRED_FEATURES = 64
ABORT_BUCKETS = 4096


class CsmvError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class Config(C.Structure):
    _fields_ = [("n_items", C.c_uint32), ("depth", C.c_uint32), ("atr_entries", C.c_uint32),
                ("client_waves", C.c_uint32), ("wg_waves", C.c_uint32), ("variant", C.c_uint32),
                ("timeout_ms", C.c_uint32), ("max_attempts", C.c_uint32),
                ("stage_ahead", C.c_uint32), ("lds_kib", C.c_uint32),
                ("n_serial_waves", C.c_uint32), ("begin_wait", C.c_uint32),
                ("client_wg_waves", C.c_uint32), ("escalate_after", C.c_uint32),
                ("commit_parts", C.c_uint32)]


class BankSpec(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("n_txns", C.c_uint64), ("n_reads", C.c_uint32),
                ("n_reads_ro", C.c_uint32), ("pct_ro", C.c_uint32), ("dist", C.c_uint32),
                ("amount_max", C.c_uint32), ("n_reads_min", C.c_uint32), ("theta", C.c_double)]


class MemcdSpec(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("n_reqs", C.c_uint64), ("n_keys", C.c_uint64),
                ("n_sets", C.c_uint32), ("n_ways", C.c_uint32), ("pct_get", C.c_uint32),
                ("batch", C.c_uint32), ("dist", C.c_uint32), ("reserved", C.c_uint32),
                ("theta", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("commits_update", C.c_uint64), ("commits_ro", C.c_uint64),
                ("aborts_dataage", C.c_uint64), ("aborts_readwrite", C.c_uint64),
                ("aborts_prevalid", C.c_uint64), ("aborts_recordage", C.c_uint64),
                ("batches", C.c_uint64), ("final_ts", C.c_uint64), ("kernel_ms", C.c_double),
                ("error", C.c_uint32), ("client_lanes", C.c_uint32),
                ("prof_ns", C.c_double * 16), ("stamp_buckets", C.c_uint32),
                ("stamp_exact", C.c_uint32), ("aborts_unknown", C.c_uint64),
                ("lds_bytes", C.c_uint32), ("reserved", C.c_uint32),
                ("escalated", C.c_uint64)]
    PROF = ("commit_idle", "commit_busy", "recv_window", "recv_ring", "recv_busy",
            "client_reply", "client_round", "client_rounds", "commit_read", "commit_resolve",
            "commit_chain", "commit_finish", "commit_spin", "commit_adv", "commit_misses")

    def as_dict(self) -> dict:
        d = {f: getattr(self, f) for f, _ in self._fields_ if f != "prof_ns"}
        d["prof"] = {k: self.prof_ns[i] for i, k in enumerate(self.PROF)}
        commits = d["commits_update"] + d["commits_ro"]
        aborts = (d["aborts_dataage"] + d["aborts_readwrite"] + d["aborts_prevalid"]
                  + d["aborts_recordage"] + d["aborts_unknown"])
        d["commits"] = commits
        d["aborts"] = aborts
        d["abort_rate"] = aborts / (aborts + commits) if aborts + commits else 0.0
        d["txn_per_s"] = commits / (d["kernel_ms"] * 1e-3) if d["kernel_ms"] > 0 else 0.0
        return d


class BankBuffers(C.Structure):
    """csmv_bank_buffers: device pointers of an allocation-free bank run."""
    _fields_ = [("init", C.c_void_p), ("final_state", C.c_void_p), ("op", C.c_void_p),
                ("addr", C.c_void_p), ("n_txns", C.c_uint64), ("log_ts", C.c_void_p),
                ("log_v", C.c_void_p), ("log_ab", C.c_void_p)]


class LogHeader(C.Structure):
    """csmv_log_header: the 64-byte header of a log file."""
    _fields_ = [("magic", C.c_char * 8), ("version", C.c_uint32), ("kind", C.c_uint32),
                ("n", C.c_uint64), ("rows", C.c_uint32), ("buckets", C.c_uint32),
                ("seed", C.c_uint64), ("checksum", C.c_uint64), ("reserved", C.c_uint8 * 16)]


LOG_BANK, LOG_CACHE, LOG_ABORTS = 1, 2, 3
ABORT_CLASSES = ("commits_update", "commits_ro", "aborts_dataage", "aborts_readwrite",
                 "aborts_prevalid", "aborts_recordage", "aborts_unknown", "escalated")


class ServeSpec(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("offered_rps", C.c_double), ("pct_cache", C.c_uint32),
                ("reserved", C.c_uint32), ("pool_budget", C.c_uint64)]


class ServeStats(C.Structure):
    _fields_ = [("cache", Stats), ("bank", Stats), ("kernel_ms", C.c_double),
                ("admit_stalls", C.c_uint64), ("peak_inflight", C.c_uint64),
                ("final_inflight", C.c_uint64), ("wallclock_khz", C.c_uint32),
                ("reserved", C.c_uint32)]


_lib: Optional[C.CDLL] = None


def lib() -> C.CDLL:
    """Loads libcsmv.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    P, U32, U64, I32, SZ = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int32, C.c_size_t
    sigs = {
        "csmv_status_str": ([C.c_int], C.c_char_p),
        "csmv_last_error": ([P], C.c_char_p),
        "csmv_create": ([C.POINTER(Config), C.POINTER(P)], C.c_int),
        "csmv_destroy": ([P], C.c_int),
        "csmv_load_i32": ([P, P, SZ], C.c_int),
        "csmv_fill_i32": ([P, I32], C.c_int),
        "csmv_read_state_i32": ([P, P, SZ], C.c_int),
        "csmv_zipf_cdf": ([U32, C.c_double, P], C.c_int),
        "csmv_bank_generate": ([P, C.POINTER(BankSpec)], C.c_int),
        "csmv_bank_read_stream": ([P, P, P], C.c_int),
        "csmv_bank_run": ([P, C.POINTER(Stats)], C.c_int),
        "csmv_bank_route": ([P, P, SZ], C.c_int),
        "csmv_bank_read_log": ([P, P, P, P], C.c_int),
        "csmv_atomic_rate": ([P, U32, U32, C.POINTER(C.c_double)], C.c_int),
        "csmv_memcd_generate": ([P, C.POINTER(MemcdSpec)], C.c_int),
        "csmv_memcd_read_stream": ([P, P, P], C.c_int),
        "csmv_memcd_run": ([P, C.POINTER(Stats)], C.c_int),
        "csmv_memcd_read_log": ([P, P, P, P, P, P], C.c_int),
        "csmv_wallclock_khz": ([P, C.POINTER(U32)], C.c_int),
        "csmv_serve_run": ([P, P, C.POINTER(ServeSpec), C.POINTER(ServeStats)], C.c_int),
        "csmv_serve_run_multi": ([P, U32, P, C.POINTER(ServeSpec), C.POINTER(ServeStats)], C.c_int),
        "csmv_memcd_load": ([P, C.POINTER(MemcdSpec), P, P], C.c_int),
        "csmv_serve_read_arrivals": ([P, P, SZ], C.c_int),
        "csmv_serve_read_latency": ([P, P, SZ], C.c_int),
        "csmv_pred_init": ([P, U64], C.c_int),
        "csmv_pred_set_params": ([P, P, SZ], C.c_int),
        "csmv_pred_get_params": ([P, P, SZ], C.c_int),
        "csmv_abort_fold": ([P, U32], C.c_int),
        "csmv_abort_read_counters": ([P, P, SZ], C.c_int),
        "csmv_abort_load_counters": ([P, P, SZ], C.c_int),
        "csmv_pred_collect": ([P, U64, C.POINTER(U64)], C.c_int),
        "csmv_pred_read_dataset": ([P, P, P, SZ], C.c_int),
        "csmv_pred_clear_dataset": ([P], C.c_int),
        "csmv_pred_train": ([P, U32, C.c_float, P], C.c_int),
        "csmv_pred_forward": ([P, P, SZ, P], C.c_int),
        "csmv_pred_grad": ([P, P, P, SZ, P, P], C.c_int),
        "csmv_pred_infer": ([P, C.c_float, P, SZ], C.c_int),
        "csmv_bank_route_learned": ([P, C.c_float, C.POINTER(U64)], C.c_int),
        "csmv_bank_read_route": ([P, P, P, C.POINTER(U64), C.POINTER(U64)], C.c_int),
        "csmv_bank_run_noalloc": ([C.POINTER(Config), C.POINTER(BankBuffers), P,
                                   C.POINTER(Stats)], C.c_int),
        "csmv_log_write_bank": ([C.c_char_p, U64, U64, U32, P, P, P], C.c_int),
        "csmv_log_read_bank": ([C.c_char_p, C.POINTER(LogHeader), P, P, P], C.c_int),
        "csmv_log_write_cache": ([C.c_char_p, U64, U64, P, P, P, P], C.c_int),
        "csmv_log_read_cache": ([C.c_char_p, C.POINTER(LogHeader), P, P, P, P], C.c_int),
        "csmv_log_write_aborts": ([C.c_char_p, U64, C.POINTER(Stats), U64, P, P, U32], C.c_int),
        "csmv_log_read_aborts": ([C.c_char_p, C.POINTER(LogHeader), P, P, P], C.c_int),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _ptr(a: np.ndarray) -> C.c_void_p:
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


def zipf_cdf(n: int, theta: float) -> np.ndarray:
    """Host Zipf table (needs no device): cdf[i] = floor(2^31 H(i+1)/H(n)), cdf[n-1] = 2^31."""
    out = np.empty(n, dtype=np.uint32)
    st = lib().csmv_zipf_cdf(n, theta, _ptr(out))
    if st:
        raise CsmvError(st, "csmv_zipf_cdf")
    return out


@dataclasses.dataclass
class Bank:
    """The bank workload (bank/csmv_sh/exe.cu:147-207): transfers and audits."""
    n_txns: int
    seed: int = 1
    n_reads: int = 2
    n_reads_ro: int = 32
    pct_ro: int = 10
    zipf: bool = False
    theta: float = 0.99
    amount_max: int = 10
    n_reads_min: int = 0     # > 0: transfers read uniform n_reads_min..n_reads accounts

    def spec(self) -> BankSpec:
        return BankSpec(self.seed, self.n_txns, self.n_reads, self.n_reads_ro, self.pct_ro,
                        1 if self.zipf else 0, self.amount_max, self.n_reads_min, self.theta)

    @property
    def rmax(self) -> int:
        return max(self.n_reads, self.n_reads_ro)


@dataclasses.dataclass
class Memcd:
    """The memcached-style cache workload (memcachedGPU/memcdSH.cu:299-448)."""
    n_reqs: int
    seed: int = 1
    n_keys: int = 10_000_000
    n_sets: int = 1 << 17
    n_ways: int = 8
    pct_get: int = 95
    batch: int = 65536
    zipf: bool = True
    theta: float = 0.99

    def spec(self) -> MemcdSpec:
        return MemcdSpec(self.seed, self.n_reqs, self.n_keys, self.n_sets, self.n_ways,
                         self.pct_get, self.batch, 1 if self.zipf else 0, 0, self.theta)

    @property
    def n_items(self) -> int:
        return 12 * self.n_sets * self.n_ways


def config(n_items: int, depth: int = 4, atr_entries: int = 4096, client_waves: int = 64,
           wg_waves: int = 8, variant="stamp", timeout_ms: int = 4000, max_attempts: int = 0,
           stage_ahead: int = 0, lds_kib: int = 0, n_serial_waves: int = 0,
           begin_wait: bool = True, client_wg_waves: int = 0, escalate_after: int = 0,
           commit_parts: int = 0) -> Config:
    """A csmv_config (0 fields take the engine's defaults)."""
    if isinstance(variant, str):
        variant = PRESETS[variant]
    return Config(n_items, depth, atr_entries, client_waves, wg_waves, variant, timeout_ms,
                  max_attempts, stage_ahead, lds_kib, n_serial_waves, 1 if begin_wait else 0,
                  client_wg_waves, escalate_after, commit_parts)


def bank_run_noalloc(cfg: Config, buffers: BankBuffers, stream: Optional[int] = None) -> dict:
    """csmv_bank_run_noalloc: a bank run on caller-owned device buffers, with no device
    allocation (the engine's fixed scratch array holds everything else)."""
    st = Stats()
    rc = lib().csmv_bank_run_noalloc(C.byref(cfg), C.byref(buffers),
                                     C.c_void_p(stream) if stream else None, C.byref(st))
    if rc:
        raise CsmvError(rc, "csmv_bank_run_noalloc")
    return st.as_dict()


class Engine:
    def __init__(self, n_items: int, depth: int = 4, atr_entries: int = 4096,
                 client_waves: int = 64, wg_waves: int = 8, variant="stamp",
                 timeout_ms: int = 4000, max_attempts: int = 0, stage_ahead: int = 0,
                 lds_kib: int = 0, n_serial_waves: int = 0, begin_wait: bool = True,
                 client_wg_waves: int = 0, escalate_after: int = 0, commit_parts: int = 0):
        self.cfg = config(n_items, depth, atr_entries, client_waves, wg_waves, variant,
                          timeout_ms, max_attempts, stage_ahead, lds_kib, n_serial_waves,
                          begin_wait, client_wg_waves, escalate_after, commit_parts)
        self.n_items = n_items
        self._h = C.c_void_p()
        st = lib().csmv_create(C.byref(self.cfg), C.byref(self._h))
        if st:
            raise CsmvError(st, "csmv_create")
        self.bank: Optional[Bank] = None

    def _chk(self, st: int, what: str):
        if st:
            msg = lib().csmv_last_error(self._h) or b""
            raise CsmvError(st, f"{what}: {msg.decode(errors='replace')}")

    def close(self):
        if self._h:
            lib().csmv_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- pool ---------------------------------------------------------------------------
    def fill(self, value: int):
        self._chk(lib().csmv_fill_i32(self._h, int(value)), "fill")

    def load(self, values: np.ndarray):
        v = np.ascontiguousarray(values, dtype=np.int32)
        self._chk(lib().csmv_load_i32(self._h, _ptr(v), v.size), "load")

    def read_state(self) -> np.ndarray:
        out = np.empty(self.n_items, dtype=np.int32)
        self._chk(lib().csmv_read_state_i32(self._h, _ptr(out), out.size), "read_state")
        return out

    # ---- bank ---------------------------------------------------------------------------
    def bank_generate(self, bank: Bank):
        self._chk(lib().csmv_bank_generate(self._h, C.byref(bank.spec())), "bank_generate")
        self.bank = bank

    def bank_read_stream(self):
        b = self.bank
        op = np.empty(b.n_txns, dtype=np.uint32)
        addr = np.empty(b.rmax * b.n_txns, dtype=np.uint32)
        self._chk(lib().csmv_bank_read_stream(self._h, _ptr(op), _ptr(addr)), "read_stream")
        return op, addr

    def bank_route(self, hot: Optional[np.ndarray]):
        """Sends transfers with hot[k] set to the serialized lane (None: no routing)."""
        if hot is None:
            self._chk(lib().csmv_bank_route(self._h, None, 0), "bank_route")
            return
        h = np.ascontiguousarray(hot, dtype=np.uint8)
        self._chk(lib().csmv_bank_route(self._h, _ptr(h), h.size), "bank_route")

    def bank_run(self) -> dict:
        st = Stats()
        self._chk(lib().csmv_bank_run(self._h, C.byref(st)), "bank_run")
        return st.as_dict()

    def bank_read_log(self):
        b = self.bank
        ts = np.empty(b.n_txns, dtype=np.uint32)
        v = np.empty(b.n_reads * b.n_txns, dtype=np.int32)
        ab = np.empty(b.n_txns, dtype=np.uint32)
        self._chk(lib().csmv_bank_read_log(self._h, _ptr(ts), _ptr(v), _ptr(ab)), "read_log")
        return ts, v, ab

    # ---- memcached-style cache ----------------------------------------------------------
    def memcd_generate(self, m: Memcd):
        self._chk(lib().csmv_memcd_generate(self._h, C.byref(m.spec())), "memcd_generate")
        self.memcd = m

    def memcd_load(self, m: Memcd, key: np.ndarray, op: np.ndarray):
        """A request stream from host arrays (one partition of a split stream); m gives the
        engine's sets, ways and batch, and n_reqs = key.size."""
        k = np.ascontiguousarray(key, dtype=np.uint64)
        o = np.ascontiguousarray(op, dtype=np.uint32)
        assert k.size == o.size
        m = dataclasses.replace(m, n_reqs=int(k.size))
        self._chk(lib().csmv_memcd_load(self._h, C.byref(m.spec()), _ptr(k), _ptr(o)), "memcd_load")
        self.memcd = m

    def memcd_read_stream(self):
        m = self.memcd
        key = np.empty(m.n_reqs, dtype=np.uint64)
        op = np.empty(m.n_reqs, dtype=np.uint32)
        self._chk(lib().csmv_memcd_read_stream(self._h, _ptr(key), _ptr(op)), "memcd_read_stream")
        return key, op

    def memcd_run(self) -> dict:
        st = Stats()
        self._chk(lib().csmv_memcd_run(self._h, C.byref(st)), "memcd_run")
        return st.as_dict()

    def memcd_read_log(self):
        n = self.memcd.n_reqs
        out = [np.empty(n, dtype=np.uint32) for _ in range(5)]
        self._chk(lib().csmv_memcd_read_log(self._h, *[_ptr(a) for a in out]), "memcd_read_log")
        return dict(zip(("ts", "res", "vsum", "aborts", "lat_ticks"), out))

    def wallclock_khz(self) -> int:
        k = C.c_uint32()
        self._chk(lib().csmv_wallclock_khz(self._h, C.byref(k)), "wallclock_khz")
        return k.value

    def atomic_rate(self, waves: int = 4096, iters: int = 256) -> float:
        r = C.c_double()
        self._chk(lib().csmv_atomic_rate(self._h, waves, iters, C.byref(r)), "atomic_rate")
        return r.value

    # ---- learned contention predictor ---------------------------------------------------
    def pred_init(self, seed: int = 1):
        self._chk(lib().csmv_pred_init(self._h, seed), "pred_init")

    def pred_set_params(self, params: np.ndarray):
        a = np.ascontiguousarray(params, dtype=np.float32)
        self._chk(lib().csmv_pred_set_params(self._h, _ptr(a), a.size), "pred_set_params")

    def pred_get_params(self) -> np.ndarray:
        out = np.empty(PRED_PARAMS, dtype=np.float32)
        self._chk(lib().csmv_pred_get_params(self._h, _ptr(out), out.size), "pred_get_params")
        return out

    def abort_fold(self, decay_shift: int = 1):
        """counters = (counters >> decay_shift) + the last run's abort log."""
        self._chk(lib().csmv_abort_fold(self._h, decay_shift), "abort_fold")

    def abort_counters(self) -> np.ndarray:
        out = np.empty(ABORT_BUCKETS, dtype=np.uint32)
        self._chk(lib().csmv_abort_read_counters(self._h, _ptr(out), out.size), "abort_counters")
        return out

    def abort_load_counters(self, counters: np.ndarray):
        a = np.ascontiguousarray(counters, dtype=np.uint32)
        self._chk(lib().csmv_abort_load_counters(self._h, _ptr(a), a.size), "abort_load")

    def pred_collect(self, max_samples: int = 1 << 62) -> int:
        n = C.c_uint64()
        self._chk(lib().csmv_pred_collect(self._h, max_samples, C.byref(n)), "pred_collect")
        return n.value

    def pred_dataset(self, n: int):
        X = np.empty((n, PRED_FEATURES), dtype=np.uint16)
        y = np.empty(n, dtype=np.float32)
        self._chk(lib().csmv_pred_read_dataset(self._h, _ptr(X), _ptr(y), n), "pred_dataset")
        return X, y

    def pred_clear_dataset(self):
        self._chk(lib().csmv_pred_clear_dataset(self._h), "pred_clear_dataset")

    def pred_train(self, steps: int, lr: float = 1e-2) -> np.ndarray:
        loss = np.empty(steps, dtype=np.float32)
        self._chk(lib().csmv_pred_train(self._h, steps, lr, _ptr(loss)), "pred_train")
        return loss

    def pred_forward(self, X: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(X, dtype=np.uint16)
        n = x.shape[0]
        p = np.empty(n, dtype=np.float32)
        self._chk(lib().csmv_pred_forward(self._h, _ptr(x), n, _ptr(p)), "pred_forward")
        return p

    def pred_grad(self, X: np.ndarray, y: np.ndarray):
        x = np.ascontiguousarray(X, dtype=np.uint16)
        yy = np.ascontiguousarray(y, dtype=np.float32)
        g = np.empty(PRED_PARAMS, dtype=np.float32)
        loss = C.c_float()
        self._chk(lib().csmv_pred_grad(self._h, _ptr(x), _ptr(yy), x.shape[0], _ptr(g),
                                       C.byref(loss)), "pred_grad")
        return g, loss.value

    def pred_infer(self, tau: float = 0.5) -> np.ndarray:
        p = np.empty(self.bank.n_txns, dtype=np.float32)
        self._chk(lib().csmv_pred_infer(self._h, tau, _ptr(p), p.size), "pred_infer")
        return p

    def bank_route_learned(self, tau: float = 0.5) -> int:
        """Routes transfers with predicted abort probability above tau to the serialized
        lane; returns how many."""
        n = C.c_uint64()
        self._chk(lib().csmv_bank_route_learned(self._h, tau, C.byref(n)), "route_learned")
        return n.value

    def bank_read_route(self):
        nn, nh = C.c_uint64(), C.c_uint64()
        self._chk(lib().csmv_bank_read_route(self._h, None, None, C.byref(nn), C.byref(nh)),
                  "read_route")
        norm = np.empty(nn.value, dtype=np.uint32)
        hot = np.empty(nh.value, dtype=np.uint32)
        self._chk(lib().csmv_bank_read_route(self._h, _ptr(norm), _ptr(hot), None, None),
                  "read_route")
        return norm, hot

    # ---- serving loop -------------------------------------------------------------------
    def _served_len(self) -> int:
        served = getattr(self, "_served", None)
        if served == "cache":
            return self.memcd.n_reqs
        if served == "bank":
            return self.bank.n_txns
        raise RuntimeError("engine has not served a stream")

    def serve_arrivals(self) -> np.ndarray:
        n = self._served_len()
        out = np.empty(n, dtype=np.uint32)
        self._chk(lib().csmv_serve_read_arrivals(self._h, _ptr(out), n), "serve_arrivals")
        return out

    def serve_latency(self) -> np.ndarray:
        """Latency of each request from arrival to commit, in wall-clock ticks."""
        n = self._served_len()
        out = np.empty(n, dtype=np.uint32)
        self._chk(lib().csmv_serve_read_latency(self._h, _ptr(out), n), "serve_latency")
        return out


def memcd_set_of(key: np.ndarray, n_sets: int) -> np.ndarray:
    """The cache set of each key (the device's mc_set_of: a splitmix64 finaliser)."""
    z = np.asarray(key, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    z ^= z >> np.uint64(31)
    return (z & np.uint64(n_sets - 1)).astype(np.uint32)


def memcd_partition(key: np.ndarray, op: np.ndarray, n_sets: int, n_parts: int):
    """Splits a cache request stream over n_parts partitions owning contiguous ranges of
    n_sets / n_parts sets (partition = set // (n_sets / n_parts)).  Within a partition a
    request's set is set mod (n_sets / n_parts), the set a cache of n_sets / n_parts sets
    gives its key, so each partition is an ordinary cache stream (csmv_memcd_load)."""
    assert n_parts >= 1 and n_sets % n_parts == 0 and (n_sets // n_parts) & (n_sets // n_parts - 1) == 0
    per = n_sets // n_parts
    part = memcd_set_of(key, n_sets) // np.uint32(per)
    return [(np.ascontiguousarray(key[part == p]), np.ascontiguousarray(op[part == p]))
            for p in range(n_parts)]


def serve(cache, bank: Optional[Engine], offered_rps: float,
          pct_cache: int = 80, seed: int = 1, pool_budget: int = 0) -> dict:
    """Runs the persistent serving loop over a cache partition (an Engine, a list of
    Engines holding disjoint ranges of sets, or None) and a bank partition (or None): a
    class that is None is not served, and the other keeps the arrival times it has in the
    mix."""
    sp = ServeSpec(seed, float(offered_rps), pct_cache, 0, pool_budget)
    st = ServeStats()
    caches = [] if cache is None else (list(cache) if isinstance(cache, (list, tuple)) else [cache])
    h_b = bank._h if bank is not None else None
    if len(caches) <= 1:
        h_c = caches[0]._h if caches else None
        rc = lib().csmv_serve_run(h_c, h_b, C.byref(sp), C.byref(st))
    else:
        arr = (C.c_void_p * len(caches))(*[e._h.value for e in caches])
        rc = lib().csmv_serve_run_multi(arr, len(caches), h_b, C.byref(sp), C.byref(st))
    for e in caches:
        e._served = "cache"
    if bank is not None:
        bank._served = "bank"
    if rc:
        e = caches[0] if caches else bank
        msg = lib().csmv_last_error(e._h) or b""
        raise CsmvError(rc, f"serve: {msg.decode(errors='replace')}")
    return {"cache": st.cache.as_dict() if caches else None,
            "bank": st.bank.as_dict() if bank is not None else None,
            "kernel_ms": st.kernel_ms, "admit_stalls": st.admit_stalls,
            "peak_inflight": st.peak_inflight, "final_inflight": st.final_inflight,
            "wallclock_khz": st.wallclock_khz}


# ---- log files (csmv.h: 64-byte header + little-endian arrays) --------------------------------
def _chk_io(rc: int, what: str):
    if rc:
        raise CsmvError(rc, what)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def log_write_bank(path: str, ts, v, aborts, seed: int = 0):
    ts, ab = _u32(ts), _u32(aborts)
    v = np.ascontiguousarray(v, dtype=np.int32)
    n = ts.size
    assert v.size % n == 0 and ab.size == n
    _chk_io(lib().csmv_log_write_bank(os.fsencode(path), seed, n, v.size // n, _ptr(ts), _ptr(v),
                                      _ptr(ab)), "log_write_bank")


def log_read_header(path: str, kind: int) -> LogHeader:
    h = LogHeader()
    fn = {LOG_BANK: lib().csmv_log_read_bank, LOG_CACHE: lib().csmv_log_read_cache,
          LOG_ABORTS: lib().csmv_log_read_aborts}[kind]
    nulls = [None] * (4 if kind == LOG_CACHE else 3)
    _chk_io(fn(os.fsencode(path), C.byref(h), *nulls), "log_read_header")
    return h


def log_read_bank(path: str):
    """Returns (header, ts, v, aborts) of a bank commit log file."""
    h = log_read_header(path, LOG_BANK)
    ts = np.empty(h.n, dtype=np.uint32)
    v = np.empty(h.rows * h.n, dtype=np.int32)
    ab = np.empty(h.n, dtype=np.uint32)
    _chk_io(lib().csmv_log_read_bank(os.fsencode(path), None, _ptr(ts), _ptr(v), _ptr(ab)),
            "log_read_bank")
    return h, ts, v, ab


def log_write_cache(path: str, log: dict, seed: int = 0):
    a = [_u32(log[k]) for k in ("ts", "res", "vsum", "aborts")]
    _chk_io(lib().csmv_log_write_cache(os.fsencode(path), seed, a[0].size, *[_ptr(x) for x in a]),
            "log_write_cache")


def log_read_cache(path: str):
    h = log_read_header(path, LOG_CACHE)
    out = {k: np.empty(h.n, dtype=np.uint32) for k in ("ts", "res", "vsum", "aborts")}
    _chk_io(lib().csmv_log_read_cache(os.fsencode(path), None,
                                      *[_ptr(out[k]) for k in ("ts", "res", "vsum", "aborts")]),
            "log_read_cache")
    return h, out


def log_write_aborts(path: str, stats: dict, attempts, counters=None, seed: int = 0):
    st = Stats()
    for k in ABORT_CLASSES:
        setattr(st, k, int(stats.get(k, 0)))
    att = _u32(attempts)
    cnt = None if counters is None else _u32(counters)
    _chk_io(lib().csmv_log_write_aborts(os.fsencode(path), seed, C.byref(st), att.size, _ptr(att),
                                        None if cnt is None else _ptr(cnt),
                                        0 if cnt is None else cnt.size), "log_write_aborts")


def log_read_aborts(path: str):
    """Returns (header, classes dict, attempts, counters or None) of an abort log file."""
    h = log_read_header(path, LOG_ABORTS)
    cl = np.empty(8, dtype=np.uint64)
    att = np.empty(h.n, dtype=np.uint32)
    cnt = np.empty(h.buckets, dtype=np.uint32) if h.buckets else None
    _chk_io(lib().csmv_log_read_aborts(os.fsencode(path), None, _ptr(cl), _ptr(att),
                                       None if cnt is None else _ptr(cnt)), "log_read_aborts")
    return h, dict(zip(ABORT_CLASSES, (int(x) for x in cl))), att, cnt
