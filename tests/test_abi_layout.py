"""The C ABI's struct layouts as gcc lays them out (tests/abi_layout.c) equal the ctypes mirror in
flashraw_amd/gpu.py field by field; the GPU-off switch and the cancel token work without a device."""
import ctypes
import os
import subprocess
import sys

import pytest

import flashraw_amd as F
from flashraw_amd import gpu as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

MIRRORS = {
    "frh_rect": G._Rect, "frh_frame": G._Frame, "frh_frame_info": G._FrameInfo, "frh_edge_spec": G._EdgeSpec,
    "frh_curve": G._Curve, "frh_tone": G._Tone, "frh_develop_spec": G._DevelopSpec, "frh_timings": G._Timings,
    "frh_dn_spec": G._DnSpec, "frh_nr_plane": G._NrPlane, "frh_nr_spec": G._NrSpec,
    "frh_server_config": G._ServerConfig, "frh_job_result": G._JobResult, "frh_server_stats": G._ServerStats,
}


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("abi") / "abi_layout")
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "abi_layout.c"), "-o", exe], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    layout = {}
    for line in out.splitlines():
        struct, field, off, size = line.split()
        layout.setdefault(struct, []).append((field, int(off), int(size)))
    return layout


def test_every_struct_is_mirrored(c_layout):
    assert set(c_layout) == set(MIRRORS)


@pytest.mark.parametrize("name", sorted(MIRRORS))
def test_ctypes_mirror_matches_c(c_layout, name):
    cls = MIRRORS[name]
    rows = c_layout[name]
    assert rows[0][0] == "-" and rows[0][2] == ctypes.sizeof(cls), (name, rows[0], ctypes.sizeof(cls))
    fields = rows[1:]
    assert [f for f, _, _ in fields] == [f for f, _ in cls._fields_], name
    for f, off, size in fields:
        d = getattr(cls, f)
        assert (d.offset, d.size) == (off, size), (name, f, d.offset, d.size, off, size)


def _in_subprocess(code, env_value):
    env = dict(os.environ)
    if env_value is None:
        env.pop("FLASHRAW_GPU", None)
    else:
        env["FLASHRAW_GPU"] = env_value
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout.strip()


@pytest.mark.parametrize("value", ["0", "off", "OFF", "false", "None", "cpu"])
def test_gpu_off_switch(value):
    # disabled_by_env (crates/flashraw-gpu/src/context.rs:734-747): available() is false and the
    # context refuses with GpuUnavailable, naming the switch, before any device call
    code = ("import ctypes, flashraw_amd as F\n"
            "L = F.lib(); h = ctypes.c_void_p()\n"
            "rc = L.frh_context_create(0, ctypes.byref(h))\n"
            "print(L.frh_available(), rc, L.frh_last_error_message().decode())\n")
    out = _in_subprocess(code, value)
    avail, rc, msg = out.split(" ", 2)
    assert avail == "0" and rc == "6" and "FLASHRAW_GPU" in msg, out


@pytest.mark.parametrize("value", [None, "1", "on", "hip"])
def test_gpu_switch_other_values_do_not_disable(value):
    code = ("import ctypes, flashraw_amd as F\n"
            "L = F.lib(); h = ctypes.c_void_p()\n"
            "rc = L.frh_context_create(0, ctypes.byref(h))\n"
            "print(rc, L.frh_last_error_message().decode())\n"
            "if rc == 0: L.frh_context_destroy(h)\n")
    out = _in_subprocess(code, value)
    assert "FLASHRAW_GPU" not in out, out


def test_cancel_token_flag_and_null():
    L = F.lib()
    t = F.CancelToken()
    assert not t.is_cancelled()
    t.cancel()
    t.cancel()  # any number of times
    assert t.is_cancelled()
    # NULL: never cancelled, and cancelling it does nothing (flashraw.h:1251-1265)
    L.frh_cancel_token_cancel(None)
    assert L.frh_cancel_token_is_cancelled(None) == 0
    L.frh_cancel_token_free(None)


def test_cancel_token_poll_cancels_for_good():
    # CancelToken::with_poll (crates/flashraw-core/src/cancel.rs:40-63): asked while not cancelled,
    # a true answer sets the flag for good, and it is not asked again
    calls = []

    def poll():
        calls.append(1)
        return len(calls) >= 3

    t = F.CancelToken(poll)
    assert [t.is_cancelled() for _ in range(5)] == [False, False, True, True, True]
    assert len(calls) == 3


def test_cancel_token_from_another_thread():
    import threading
    t = F.CancelToken()
    th = threading.Thread(target=t.cancel)
    th.start()
    th.join()
    assert t.is_cancelled()
