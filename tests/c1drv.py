"""ctypes binding of the external bank driver (tests/driver/libc1drv.so), a kernel written
only against include/ that runs transactions on an engine through csmv_kernel_prepare /
csmv_kernel_finish."""
import ctypes as C
import os

import numpy as np

import csmv

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "driver", "libc1drv.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        csmv.lib()
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.c1drv_run.argtypes = [P, P, P, C.c_uint64, C.c_uint32, P, P, P, C.POINTER(csmv.Stats)]
        L.c1drv_run.restype = C.c_int
        L.lanechk_run.argtypes = [P, P, P, C.c_uint32, P, P, P]
        L.lanechk_run.restype = C.c_int
        _lib = L
    return _lib


def run(engine, op: np.ndarray, addr: np.ndarray, rmax: int):
    """Runs the stream through the external driver; returns (stats, ts, v, aborts)."""
    n = op.size
    op = np.ascontiguousarray(op, dtype=np.uint32)
    addr = np.ascontiguousarray(addr, dtype=np.uint32)
    assert addr.size == rmax * n
    ts = np.zeros(n, dtype=np.uint32)
    v = np.zeros(2 * n, dtype=np.int32)
    ab = np.zeros(n, dtype=np.uint32)
    st = csmv.Stats()
    rc = lib().c1drv_run(engine._h, op.ctypes.data, addr.ctypes.data, n, rmax, ts.ctypes.data,
                         v.ctypes.data, ab.ctypes.data, C.byref(st))
    if rc:
        engine._chk(rc, "c1drv_run")
    return st.as_dict(), ts, v, ab


def lane_check(v: np.ndarray, v64: np.ndarray, src: np.ndarray):
    """Runs the lane-primitive kernel on waves of 64 words; returns (inclusive scans,
    gathers v[src], reductions [min, max, sum, or64 lo, or64 hi] per wave)."""
    v = np.ascontiguousarray(v, dtype=np.uint32)
    v64 = np.ascontiguousarray(v64, dtype=np.uint64)
    src = np.ascontiguousarray(src, dtype=np.uint32)
    assert v.size % 64 == 0 and v.size == v64.size == src.size
    nw = v.size // 64
    scan = np.zeros_like(v)
    gather = np.zeros_like(v)
    red = np.zeros(5 * nw, dtype=np.uint32)
    rc = lib().lanechk_run(v.ctypes.data, v64.ctypes.data, src.ctypes.data, nw, scan.ctypes.data,
                           gather.ctypes.data, red.ctypes.data)
    if rc:
        raise RuntimeError(f"lanechk_run: HIP error {rc}")
    return scan, gather, red.reshape(nw, 5)
