"""The C ABI library loads without a GPU and exports every symbol include/flashraw_hip.h declares;
host-side validation and tables work without a device."""
import ctypes
import os
import re

import numpy as np

import flashraw_amd as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "flashraw_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(frh_[a-z0-9_]+)\s*\(", src)))


def test_exports_every_declared_symbol():
    L = F.lib()
    names = _declared()
    assert len(names) >= 18
    for n in names:
        assert hasattr(L, n), n


def test_srgb8_table_matches_fixture_and_endpoints():
    t = F.srgb8_table()
    fixture = np.fromfile(os.path.join(ROOT, "tests", "golden", "srgb8_table.bin"), dtype=np.uint8)
    assert np.array_equal(t, fixture)
    # srgb_table_endpoints (crates/flashraw-gpu/src/kernels.rs:620-627)
    assert t[0] == 0 and t[4096] == 255 and np.all(np.diff(t.astype(int)) >= 0)


def test_no_device_is_reported_not_hidden():
    if F.available():
        return
    h = ctypes.c_void_p()
    rc = F.lib().frh_context_create(0, ctypes.byref(h))
    assert rc == 6  # GpuUnavailable
    assert b"device" in F.lib().frh_last_error_message()
