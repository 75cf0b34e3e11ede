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
    for s in range(8):
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
