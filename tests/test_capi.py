"""The C-ABI library loads, exports every symbol include/csmv.h declares, and its
host-only entry points work without a device."""
import ctypes
import os
import re

import numpy as np
import pytest

import csmv
import orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "csmv.h")).read()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(csmv_\w+)\s*\(", text, re.M)))


def test_exports_every_declared_symbol():
    lib = csmv.lib()
    names = declared_functions()
    assert len(names) >= 13
    for n in names:
        assert hasattr(lib, n), n


def test_status_strings():
    lib = csmv.lib()
    for s in range(9):
        assert lib.csmv_status_str(s)
    assert lib.csmv_status_str(0) == b"ok"


def test_host_zipf_table_is_the_oracles():
    for n, th in ((1000, 0.99), (1 << 16, 0.99), (10, 0.0), (777, 1.2)):
        assert np.array_equal(csmv.zipf_cdf(n, th), orc.zipf_cdf(n, th))


def test_create_without_device_reports_no_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is visible")
    with pytest.raises(csmv.CsmvError) as ei:
        csmv.Engine(4096)
    assert ei.value.status == 5


def test_bad_config_is_rejected():
    with pytest.raises(csmv.CsmvError) as ei:
        csmv.Engine(4096, atr_entries=1000)   # not a power of two
    assert ei.value.status in (2, 5)


def test_escalation_needs_the_stamp_preset():
    """Escalated transactions are executed on the serialized lane, which reads the stamp
    variant's committed-value tables: another variant with escalation is rejected before
    any device call."""
    for v in ("vanilla", "batchedAtrInsert", "clientWriteback", "collabValidation"):
        with pytest.raises(csmv.CsmvError) as ei:
            csmv.Engine(4096, variant=v, escalate_after=2)
        assert ei.value.status == 2


STRUCTS = {  # ctypes mirror -> C type in include/csmv.h
    "Config": "csmv_config",
    "BankSpec": "csmv_bank_spec",
    "MemcdSpec": "csmv_memcd_spec",
    "Stats": "csmv_stats",
    "ServeSpec": "csmv_serve_spec",
    "ServeStats": "csmv_serve_stats",
    "BankBuffers": "csmv_bank_buffers",
    "LogHeader": "csmv_log_header",
}


def test_struct_layouts_match_the_header(tmp_path):
    """Every ctypes mirror has the size and field offsets that the C compiler gives the
    header's struct (a field added on one side only shows up here)."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "csmv.h"', "int main(void) {"]
    for py, c in STRUCTS.items():
        lines.append(f'printf("{py} size %zu\\n", sizeof({c}));')
        for name, _ in getattr(csmv, py)._fields_:
            lines.append(f'printf("{py} {name} %zu\\n", offsetof({c}, {name}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    seen = 0
    for line in out:
        if not line:
            continue
        py, name, val = line.split()
        st = getattr(csmv, py)
        got = ctypes.sizeof(st) if name == "size" else getattr(st, name).offset
        assert got == int(val), (py, name, got, val)
        seen += 1
    assert seen == sum(1 + len(getattr(csmv, p)._fields_) for p in STRUCTS)


def test_external_driver_library_loads():
    """The external driver (built against include/ only) loads, finds libcsmv.so through
    its runpath and exports its entry point; the engine exports the prepare/finish pair
    that driver calls."""
    import c1drv
    assert hasattr(c1drv.lib(), "c1drv_run")
    for n in ("csmv_kernel_prepare", "csmv_kernel_finish"):
        assert hasattr(csmv.lib(), n)


def test_noalloc_rejects_before_touching_a_device():
    """The allocation-free bank run checks its arguments and the scratch capacity on the
    host: a null buffer is an argument error, a pool beyond the fixed array a capacity
    error, with or without a device."""
    bufs = csmv.BankBuffers(1, 1, 1, 1, 1 << 10, 1, 1, 1)
    with pytest.raises(csmv.CsmvError) as ei:
        csmv.bank_run_noalloc(csmv.config(4096), csmv.BankBuffers(0, 1, 1, 1, 1 << 10, 1, 1, 1))
    assert ei.value.status == 2
    with pytest.raises(csmv.CsmvError) as ei:
        csmv.bank_run_noalloc(csmv.config(1 << 20, depth=8), bufs)
    assert ei.value.status == 3
    with pytest.raises(csmv.CsmvError) as ei:
        csmv.bank_run_noalloc(csmv.config(4096, n_serial_waves=4, client_waves=8), bufs)
    assert ei.value.status == 2


def test_cache_partition_split_matches_the_oracle_set_hash():
    """The host split of a cache stream over partitions uses the device's set hash (pinned
    here against the oracle's), and each partition's local set is the set a cache of
    n_sets / n_parts sets gives the key."""
    import orc
    spec = orc.memcd_spec(20000, seed=3, n_keys=100_000, n_sets=1 << 10, zipf=True)
    key, op = orc.memcd_gen(spec)
    lib = orc._memcd_lib()
    want = np.array([lib.orc_memcd_set_of(int(k), 1 << 10) for k in key[:2000]], dtype=np.uint32)
    assert np.array_equal(csmv.memcd_set_of(key[:2000], 1 << 10), want)
    parts = csmv.memcd_partition(key, op, 1 << 10, 4)
    assert sum(k.size for k, _ in parts) == key.size
    for p, (k, o) in enumerate(parts):
        g = csmv.memcd_set_of(k, 1 << 10)
        assert np.all(g // 256 == p)
        assert np.array_equal(g % 256, csmv.memcd_set_of(k, 256))
